"""The combined stress configuration at test size: 640 noise-displaced
icospheres with 70% thin-film materials (the 10 M-triangle 4K run of
bench.py's stress leg uses the same generator with 15,680-triangle spheres).

CPU: the generator's material mix.
GPU: the wavefront render equals the megakernel bit for bit; the guided
render with α = 0 equals the plain one; lam_set_camera gives the image that
uploading the scene with that camera gives; the adjoint over every
thin-film thickness is finite, nonzero and linear in ∂loss/∂rgb.
"""
import numpy as np
import pytest

import lambda_amd as la


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def stress(tmp_path_factory):
    from lambda_amd import scenes
    kw = dict(seed=5, n_inst=640, freq=10, material_p=(0.15, 0.15, 0.7))
    wide = la.load_nls(scenes.icosphere_field(str(tmp_path_factory.mktemp("c5w")), width=96, **kw))
    square = la.load_nls(scenes.icosphere_field(str(tmp_path_factory.mktemp("c5s")), width=40, ar=1.0, **kw))
    return wide, square


def test_stress_scene_materials(stress):
    wide, _ = stress
    types = np.array([m.type for m in wide.materials()])
    assert len(types) >= 641  # ground + one per sphere (+ the lights' own)
    assert 0.6 * 640 < np.sum(types == la.MAT_THINFILM) < 0.8 * 640


@pytest.mark.gpu
def test_stress_render_paths_agree(ctx, stress):
    wide, square = stress
    ctx.upload(wide)
    pr = la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=2)
    a = ctx.render(pr)
    assert np.isfinite(a).all() and a.mean() > 0
    b = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=2, flags=la.LAM_FLAG_MEGAKERNEL))
    assert np.array_equal(a, b)
    ctx.guide_reset(1)
    g = ctx.render_guided(pr, la.guide_params(alpha=0.0, train_passes=0))
    assert np.array_equal(a, g)
    # another camera on the same geometry, without uploading again
    ctx.set_camera(square)
    c = ctx.render(pr)
    assert c.shape == (square.height, square.width, 3)
    ctx.upload(square)
    assert np.array_equal(c, ctx.render(pr))


@pytest.mark.gpu
def test_stress_adjoint(ctx, stress):
    _, square = stress
    ctx.upload(square)
    films = [i for i, m in enumerate(square.materials()) if m.type == la.MAT_THINFILM]
    dp = la.DiffParams(square, d=films)
    p = la.params(spp=2, max_scatterings=6, sample_p=0.8, hero=4, seed=3)
    rng = np.random.default_rng(1)
    shape = (square.height, square.width, 3)
    A = rng.normal(size=shape).astype(np.float32)
    B = rng.normal(size=shape).astype(np.float32)
    ga, gb, gab = ctx.adjoint(p, dp, A), ctx.adjoint(p, dp, B), ctx.adjoint(p, dp, A + B)
    # a 40x40 view at 2 spp reaches a fraction of the 448 films
    assert np.isfinite(gab).all() and np.count_nonzero(gab) > 20
    assert np.allclose(gab, ga + gb, rtol=1e-4, atol=1e-5 * np.abs(gab).max())
