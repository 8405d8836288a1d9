"""BVHs: the product's quantized BVH8 (conservative, complete, bounded
stack) and the oracle's restatement of the reference builder
(lib/numli/cg.hpp:632-749), pinned by the reference's own statistics."""
import ctypes as C
import gzip
import os

import numpy as np
import pytest

import lambda_amd as la
from oracle.binding import Oracle
from tests.conftest import random_rays

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def teapot(tmp_path_factory):
    d = tmp_path_factory.mktemp("teapot")
    obj = d / "teapot_10.obj"
    obj.write_bytes(gzip.decompress(open(os.path.join(GOLD, "teapot_10.obj.gz"), "rb").read()))
    text = ('{"camera":{"pos":[0,-60,12],"look_at":[0,0,12],"up":[0,0,1],"fov":30,"width":64},'
            '"lights":[{"name":"l","type":"point","pos":[0,0,30],"radiant_intensity":{"val":1}}],'
            '"objects":[{"name":"t","type":"trimesh","transform":{},"material":"m","source":"%s"}],'
            '"materials":[{"name":"m","type":"lambertian","albedo":{"val":0.5}}]}' % obj)
    return la.Scene.from_text(text)


def test_reference_builder_statistics_on_teapot(teapot):
    """The reference's builder on teapot_10 gives 8,891 nodes, 1,895 interior,
    max depth 10 (edges) and 1.87 triangles per leaf."""
    assert teapot.n_tris == 13054
    st = Oracle(teapot).bvh_stats()
    assert st["nodes"] == 8891 and st["interior"] == 1895
    assert st["depth"] == 11  # levels, i.e. 10 edges
    assert abs(teapot.n_tris / st["leaves"] - 1.87) < 0.01


def _bvh8(tri_v):
    tv = np.ascontiguousarray(tri_v, dtype=np.float32).reshape(-1)
    st, cons = la.SceneStats(), C.c_int()
    rc = la.lib().lamh_bvh8_check(tv.ctypes.data_as(C.POINTER(C.c_float)), len(tv) // 9, C.byref(st), C.byref(cons))
    assert rc == 0
    return st, cons.value


def test_bvh8_teapot(teapot):
    st, cons = _bvh8(teapot.tri_v())
    assert cons == 1 and st.stack_bound <= 64
    assert st.n_leaves * 4 >= teapot.n_tris


@pytest.mark.parametrize("kind", ["soup", "coplanar", "identical", "slivers"])
def test_bvh8_conservative_and_complete(kind):
    rng = np.random.default_rng(4)
    n = 20000
    if kind == "soup":
        c = rng.uniform(-50, 50, (n, 1, 3))
        t = c + rng.normal(scale=0.5, size=(n, 3, 3))
    elif kind == "coplanar":
        t = rng.uniform(-1, 1, (n, 3, 3))
        t[..., 2] = 0.25
    elif kind == "identical":
        t = np.tile(rng.uniform(-1, 1, (1, 3, 3)), (n, 1, 1))
    else:
        c = rng.uniform(-1e3, 1e3, (n, 1, 3))
        t = c + rng.normal(scale=1e-4, size=(n, 3, 3)) * np.array([1, 1e3, 1])
    st, cons = _bvh8(t)
    assert cons == 1, "a quantized child box does not contain its triangles"
    assert st.stack_bound <= 64, st.stack_bound


def test_field_scene_bvh8(field_scene):
    st, cons = _bvh8(field_scene.tri_v())
    assert cons == 1 and st.stack_bound <= 64


@pytest.mark.parametrize("scene_name", ["mixed_scene", "field_scene"])
def test_oracle_bvh_equals_brute_force(request, scene_name):
    scene = request.getfixturevalue(scene_name)
    o = Oracle(scene)
    rng = np.random.default_rng(2)
    v = scene.tri_v().reshape(-1, 3)
    rays = random_rays(rng, 4000, v.min(0) - 2, v.max(0) + 2)
    a, b = o.closest(rays), o.closest(rays, brute=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    rays["tmax"] = rng.uniform(0, 20, len(rays))
    assert np.array_equal(o.occluded(rays), o.occluded(rays, brute=True))


def test_oracle_teapot_bvh_equals_brute_force(teapot):
    o = Oracle(teapot)
    rng = np.random.default_rng(3)
    rays = random_rays(rng, 1500, [-4, -4, -1], [4, 4, 4])
    assert np.array_equal(o.closest(rays).view(np.uint32), o.closest(rays, brute=True).view(np.uint32))


def test_ties_go_to_lowest_object(tmp_path):
    """two instances of the same mesh at the same place: every hit is a tie
    in t, and the lower object id wins (the reference would keep the later)."""
    p = tmp_path / "q.obj"
    p.write_text("v -1 -1 0\nv 1 -1 0\nv 1 1 0\nv -1 1 0\nf 1 2 3\nf 1 3 4\n")
    objs = ",".join('{"name":"q%d","type":"trimesh","transform":{},"material":"m","source":"%s"}' % (i, p)
                    for i in range(3))
    text = ('{"camera":{"pos":[0,-5,1],"look_at":[0,0,0],"up":[0,0,1],"fov":40,"width":16},'
            '"lights":[{"name":"l","type":"point","pos":[0,0,5],"radiant_intensity":{"val":1}}],'
            '"objects":[%s],"materials":[{"name":"m","type":"lambertian","albedo":{"val":0.5}}]}' % objs)
    s = la.Scene.from_text(text)
    o = Oracle(s)
    rng = np.random.default_rng(0)
    rays = np.zeros(1000, dtype=la.RAY_DTYPE)
    rays["o"] = np.c_[rng.uniform(-0.9, 0.9, (1000, 2)), np.full(1000, 3.0)]
    rays["d"] = [0, 0, -1]
    rays["tmax"] = 3.4e38
    rays["skip"] = la.LAM_NONE
    for brute in (True, False):
        h = o.closest(rays, brute=brute)
        assert np.all(h["obj"] == 0) and np.all(h["t"] == np.float32(3.0))
    rays["skip"] = 0
    assert np.all(o.closest(rays)["obj"] == 1)


@pytest.mark.parametrize("scene_name", ["box_scene", "mixed_scene", "field_scene", "crowd_scene"])
def test_host_traversal_of_the_product_bvh_is_exact(request, scene_name):
    """lamh_trace_counts (the kernels' BVH8 traversal, quantized-frame slab
    test and all, restated on the host) finds the oracle's brute-force
    closest hits and occlusion on random rays, rays from inside the scene,
    axis-parallel rays (zero direction components) and short shadow rays:
    its culling keeps every box a ray enters"""
    scene = request.getfixturevalue(scene_name)
    o = Oracle(scene)
    rng = np.random.default_rng(21)
    v = scene.tri_v().reshape(-1, 3)
    rays = random_rays(rng, 6000, v.min(0) - 1, v.max(0) + 1)
    axis = random_rays(rng, 3000, v.min(0), v.max(0))
    d = np.zeros((3000, 3), np.float32)
    d[np.arange(3000), rng.integers(0, 3, 3000)] = rng.choice([-1.0, 1.0], 3000)
    two = rng.random(3000) < 0.5
    d[two, (np.argmax(np.abs(d[two]), axis=1) + 1) % 3] = rng.uniform(-1, 1, two.sum())
    axis["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([rays, axis])
    h, _ = la.host_trace(scene, rays)
    b = o.closest(rays, brute=True)
    assert np.array_equal(h.view(np.uint32), b.view(np.uint32))
    rays["tmax"] = rng.uniform(0, 10, len(rays)).astype(np.float32)
    occ, _ = la.host_trace(scene, rays, any_hit=True)
    assert np.array_equal(occ, o.occluded(rays, brute=True))
