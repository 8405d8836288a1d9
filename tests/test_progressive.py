"""Progressive rendering, film checkpoints and the device tone map.

GPU: a render in pieces (LAM_FLAG_ACCUMULATE) equals the one-shot render
bit for bit, also after a checkpoint is taken and restored into a fresh
context; accumulating without a matching film is an error; the device tone
map matches the host's (lamh_tone_map + rgb24) to one 8-bit step.
"""
import numpy as np
import pytest

import lambda_amd as la

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("flags", [0, la.LAM_FLAG_MEGAKERNEL, la.LAM_FLAG_NO_PIPE])
def test_pieces_equal_one_render(ctx, box_scene, flags):
    ctx.upload(box_scene)
    one = ctx.render(la.params(spp=7, max_scatterings=8, sample_p=0.8, hero=4, seed=3, spp_offset=2, flags=flags))
    p = la.params(spp=7, max_scatterings=8, sample_p=0.8, hero=4, seed=3, spp_offset=2, flags=flags)
    steps = list(ctx.render_progressive(p, chunk=3))
    assert [s for _, s, _ in steps] == [3, 6, 7]
    assert np.array_equal(steps[-1][0], one)
    assert steps[-1][2] == 0.0


def test_checkpoint_resume(ctx, box_scene):
    ctx.upload(box_scene)
    one = ctx.render(la.params(spp=6, max_scatterings=8, sample_p=0.8, hero=4, seed=9))
    ctx.render(la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=4, seed=9))
    film, spp, hero = ctx.film_get()
    assert (spp, hero) == (4, 4) and film.shape == (box_scene.height, box_scene.width, 8)
    other = la.Context(0)
    try:
        other.upload(box_scene)
        other.film_set(film, spp, hero)
        rest = other.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=9, spp_offset=4,
                                      flags=la.LAM_FLAG_ACCUMULATE))
    finally:
        other.close()
    assert np.array_equal(rest, one)


def test_accumulate_needs_a_film(ctx, box_scene):
    ctx.upload(box_scene)  # a new upload starts without a film
    with pytest.raises(la.LambdaError) as e:
        ctx.render(la.params(spp=1, max_scatterings=4, sample_p=0.8, hero=4, flags=la.LAM_FLAG_ACCUMULATE))
    assert e.value.code == la.LAM_E_INVALID
    ctx.render(la.params(spp=1, max_scatterings=4, sample_p=0.8, hero=16))
    with pytest.raises(la.LambdaError):  # hero count differs from the film's
        ctx.render(la.params(spp=1, max_scatterings=4, sample_p=0.8, hero=4, flags=la.LAM_FLAG_ACCUMULATE))


def test_device_tone_map_matches_host(ctx, box_scene):
    ctx.upload(box_scene)
    img = ctx.render(la.params(spp=8, max_scatterings=8, sample_p=0.8, hero=4, seed=2))
    dev = ctx.tone_map8(0.18)
    tm = la.tone_map(img, 0.18)
    t64 = tm.astype(np.float64)
    s = np.where(t64 > 0.0031308, 1.055 * np.power(np.maximum(t64, 0.0), 1 / 2.4) - 0.055, 12.92 * t64)
    host = np.clip(np.floor(255 * s + 0.5), 0, 255).astype(np.int32)
    d = np.abs(dev.astype(np.int32) - host)
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3, (d.max(), np.mean(d > 0))
    assert dev.mean() > 10


def test_returned_images_own_their_memory(ctx, box_scene):
    """Rendered images live in page-locked buffers that are reused once
    released: an image kept by the caller is never overwritten by a later
    render, and a released buffer comes back for the next render."""
    ctx.upload(box_scene)
    p = lambda seed: la.params(spp=2, max_scatterings=4, sample_p=0.8, hero=4, seed=seed)
    a = ctx.render(p(1))
    a_copy = a.copy()
    frames = [ctx.render(p(2 + i)) for i in range(6)]  # more than the pool keeps
    assert np.array_equal(a, a_copy)
    assert len({f.ctypes.data for f in frames} | {a.ctypes.data}) == 7
    assert not any(np.array_equal(f, a) for f in frames)
    b_ref = ctx.render(p(2))
    assert np.array_equal(b_ref, frames[0])
    del frames, b_ref
    c = ctx.render(p(1))
    assert np.array_equal(c, a)
    assert c.flags.writeable and c.dtype == np.float32 and c.shape == a.shape
