"""Parity at the benchmark workloads' full sizes (the scenes bench.py
measures, built by the same generators with the same seeds):

  icosphere field   64 x 15,680-tri instances (1,003,520 tris), 1920x1080,
                    32 bounces, p = 0.8, hero 4 (the headline render)
  film room         box room with 64 thin-film spheres, 512x512, 16 bounces
                    (the differentiation workload)
  glossy interior   200 Blinn rounded boxes + room shell (470,424 tris),
                    1024x1024, 32 bounces (the guiding workload)

Bars as in test_gpu_parity: closest hits and shadow results bit-exact
against the oracle on ray batches of the full scene; per-path spectral
radiance within 1e-4 relative for (pixel, sample) pairs spread over the
whole frame; and, for the headline frame, a same-sample 1-spp render of all
2,073,600 pixels against the FP32 oracle's, through both BVH builders.
"""
import numpy as np
import pytest

import lambda_amd as la
from lambda_amd import scenes
from oracle import binding as ob
from tests.test_gpu_parity import _assert_hits_equal, _bounce_rays, _check_paths
from tests.conftest import random_rays

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "icosphere_field": (lambda d: scenes.icosphere_field(d, seed=2), 32),
    "film_room": (lambda d: scenes.box_room_films(d, seed=3, width=512), 16),
    "glossy_interior": (lambda d: scenes.glossy_interior(d, seed=4, width=1024), 32),
}


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    cache = {}

    def get(name):
        if name not in cache:
            gen, bounces = WORKLOADS[name]
            s = la.load_nls(gen(str(tmp_path_factory.mktemp(name))))
            cache[name] = (s, ob.Oracle(s), bounces)
        return cache[name]
    return get


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_hits_bit_exact(ctx, full, name):
    scene, orc, _ = full(name)
    print(f"{name}: {scene.n_tris} tris, {scene.width}x{scene.height}", flush=True)
    ctx.upload(scene)
    rng = np.random.default_rng(21)
    v = scene.tri_v().reshape(-1, 3)
    lo, hi = v.min(axis=0), v.max(axis=0)
    rays = random_rays(rng, 1 << 17, lo, hi)
    o = orc.closest(rays)
    _assert_hits_equal(ctx.extend(rays), o, f"{name} random rays")
    b = _bounce_rays(rng, o, rays, 1 << 16)
    _assert_hits_equal(ctx.extend(b), orc.closest(b), f"{name} bounce rays")
    # shadow queries with random windows and skipped objects
    rays["tmax"] = rng.uniform(0.0, np.linalg.norm(hi - lo), len(rays)).astype(np.float32)
    rays["skip"] = np.where(rng.random(len(rays)) < 0.5,
                            rng.integers(0, scene.desc.n_objects, len(rays)), la.LAM_NONE).astype(np.uint32)
    go, oo = ctx.occluded(rays), orc.occluded(rays)
    assert np.array_equal(go, oo), f"{np.sum(go != oo)} shadow results differ"
    assert 0.02 < oo.mean() < 0.98


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_paths(ctx, full, name):
    scene, orc, bounces = full(name)
    ctx.upload(scene)
    npix = scene.width * scene.height
    rng = np.random.default_rng(22)
    pix = rng.integers(0, npix, 16384).astype(np.uint32)
    smp = rng.integers(0, 1024, 16384).astype(np.uint32)
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=9)
    g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
    assert np.isfinite(o).all() and (o > 0).mean() > 0.05
    _check_paths(g, o)


def test_full_size_frame_same_samples(ctx, full):
    """the headline frame at 1 spp: every pixel against the FP32 oracle's
    render of the same samples, with the host-built and the GPU-built BVH"""
    scene, orc, bounces = full("icosphere_field")
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=5)
    o = orc.render(pr)
    for gpu_bvh in (False, True):
        ctx.upload(scene, gpu_bvh=gpu_bvh)
        g = ctx.render(pr)
        assert g.shape == o.shape == (1080, 1920, 3)
        relmse = np.mean((g - o) ** 2) / np.mean(o ** 2)
        assert relmse < 1e-8, (gpu_bvh, relmse)
