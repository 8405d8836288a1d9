"""Edge cases of the render path: scenes with no triangles and with no
objects, meshes of 1, 2, 9, 1,023, 1,024, 1,025 and 3,000 triangles (the
host and the device BVH builds), 1x1 and 3x5 images, ray batches of 0, 1, 63, 65 and 6,145 rays,
and a render of zero samples.

CPU: the loader's image sizes for the tiny cameras.
GPU: closest hits and occlusion bit-exact against the oracle; same-sample
renders against the oracle (as for the full scenes); the wavefront render
equals the megakernel and the one-lane render bit for bit; spp = 0 is an
error (the reference's loop would divide 0 by 0, renderer.cpp:51-52).
"""
import json

import numpy as np
import pytest

import lambda_amd as la
from oracle import binding as ob
from tests.conftest import random_rays
from tests.scene_text import mixed_scene_text


def _scene_json(d):
    return json.loads(mixed_scene_text(d))


def _tiny_mesh(d, k, seed):
    """k random triangles in front of the camera of the mixed scene"""
    from lambda_amd import scenes
    rng = np.random.default_rng(seed)
    c = rng.uniform([-12, -8, 2], [12, 8, 20], (k, 1, 3))
    V = (c + rng.uniform(-4, 4, (k, 3, 3))).reshape(-1, 3)
    F = np.arange(3 * k).reshape(k, 3)
    e1, e2 = V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    path = f"{d}/tiny{k}.obj"
    scenes.write_obj(path, V, F, np.repeat(n, 3, axis=0))
    return path


def _make(d, kind):
    j = _scene_json(d)
    if kind == "analytic":
        j["objects"] = [o for o in j["objects"] if o["type"] != "trimesh"]
    elif kind == "no_objects":
        j["objects"] = []
    elif kind.startswith("tri"):
        k = int(kind[3:])
        j["objects"] = [o for o in j["objects"] if o["name"] in ("bot", "back")] + [
            {"name": "tiny", "type": "trimesh", "source": _tiny_mesh(d, k, seed=k), "transform": {},
             "material": "smooth"}]
    elif kind == "1x1":
        j["camera"] = dict(j["camera"], width=1, ar=1.0)
    elif kind == "3x5":
        j["camera"] = dict(j["camera"], width=3, ar=0.6)
    return la.Scene.from_text(json.dumps(j))


# 1,023 .. 3,000 triangles straddle the 1,024-leaf blocks of the device
# builder's bottom-up pass
KINDS = ["analytic", "no_objects", "tri1", "tri2", "tri9", "tri1023", "tri1024", "tri1025", "tri3000"]


@pytest.fixture(scope="module")
def edge_scenes(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("edge"))
    return {k: _make(d, k) for k in KINDS + ["1x1", "3x5"]}


def test_tiny_camera_sizes(edge_scenes):
    assert (edge_scenes["1x1"].width, edge_scenes["1x1"].height) == (1, 1)
    assert (edge_scenes["3x5"].width, edge_scenes["3x5"].height) == (3, 5)
    assert edge_scenes["analytic"].n_tris == 0 and edge_scenes["no_objects"].n_tris == 0
    assert [edge_scenes[k].n_tris for k in ("tri1", "tri2", "tri9", "tri1025")] == [1, 2, 9, 1025]


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


def _hits_equal(g, o):
    bits = lambda h: np.stack([h["t"].view(np.uint32), h["tri"], h["obj"], h["prim"]], axis=1)
    return np.array_equal(bits(g), bits(o))


@pytest.mark.gpu
@pytest.mark.parametrize("gpu_bvh", [False, True], ids=["host_bvh", "gpu_bvh"])
@pytest.mark.parametrize("kind", KINDS)
def test_edge_scene_hits_and_render(ctx, edge_scenes, kind, gpu_bvh):
    scene = edge_scenes[kind]
    ctx.upload(scene, gpu_bvh=gpu_bvh)
    orc = ob.Oracle(scene)
    rays = random_rays(np.random.default_rng(3), 20000, np.array([-20, -25, 0.5]), np.array([20, 25, 23.5]))
    g, o = ctx.extend(rays), orc.closest(rays)
    assert _hits_equal(g, o)
    if scene.n_tris:
        assert (o["tri"] != la.LAM_NONE).any()
    assert np.array_equal(ctx.occluded(rays), orc.occluded(rays))
    pr = la.params(spp=2, max_scatterings=6, sample_p=0.8, hero=16, seed=4)
    gi, oi = ctx.render(pr), orc.render(pr)
    assert np.isfinite(gi).all()
    if not oi.any():
        assert not gi.any()
    else:
        assert np.mean((gi - oi) ** 2) / np.mean(oi ** 2) < 1e-8
    assert np.array_equal(gi, ctx.render(la.params(spp=2, max_scatterings=6, sample_p=0.8, hero=16, seed=4,
                                                   flags=la.LAM_FLAG_MEGAKERNEL)))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["1x1", "3x5"])
def test_tiny_images(ctx, edge_scenes, kind):
    scene = edge_scenes[kind]
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    pr = la.params(spp=16, max_scatterings=8, sample_p=0.8, hero=4, seed=2)
    g, o = ctx.render(pr), orc.render(pr)
    assert g.shape == (scene.height, scene.width, 3)
    assert np.mean((g - o) ** 2) / np.mean(o ** 2) < 1e-8
    for fl in (la.LAM_FLAG_MEGAKERNEL, la.LAM_FLAG_NO_PIPE, la.LAM_FLAG_NO_GRAPH):
        assert np.array_equal(g, ctx.render(la.params(spp=16, max_scatterings=8, sample_p=0.8, hero=4, seed=2,
                                                      flags=fl))), fl


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 63, 65, 6145])
def test_ragged_ray_batches(ctx, field_scene, n):
    ctx.upload(field_scene)
    orc = ob.Oracle(field_scene)
    v = field_scene.tri_v().reshape(-1, 3)
    rays = random_rays(np.random.default_rng(n), n, v.min(axis=0), v.max(axis=0))
    g = ctx.extend(rays)
    assert len(g) == n
    assert _hits_equal(g, orc.closest(rays))
    assert np.array_equal(ctx.occluded(rays), orc.occluded(rays))


@pytest.mark.gpu
def test_zero_samples_is_an_error(ctx, box_scene):
    ctx.upload(box_scene)
    with pytest.raises(la.LambdaError, match="spp"):
        ctx.render(la.params(spp=0, max_scatterings=4, sample_p=0.8, hero=4, seed=1))
    # the context stays usable
    assert np.isfinite(ctx.render(la.params(spp=1, max_scatterings=4, sample_p=0.8, hero=4, seed=1))).all()
