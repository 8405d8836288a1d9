"""lam_extend_dev_noalloc (static device scratch, no device allocation)
against lam_extend_dev on the same device rays: bit-exact hits, on two
calls in a row (the second with a smaller batch, so stale queue counters
or spill columns of the first would show), and its size limit."""
import numpy as np
import pytest

import lambda_amd as la
from tests.conftest import random_rays

pytestmark = pytest.mark.gpu


def _bits(h):
    return np.stack([h["t"].view(np.uint32), h["tri"], h["obj"], h["prim"]], axis=1)


def test_noalloc_extend_matches_extend_dev(field_scene):
    ctx = la.Context(0)
    try:
        ctx.upload(field_scene)
        rng = np.random.default_rng(21)
        v = field_scene.tri_v().reshape(-1, 3)
        for n in (300000, 4099):
            rays = random_rays(rng, n, v.min(axis=0), v.max(axis=0))
            d_rays = la.DeviceBuffer.from_array(rays)
            d_a = la.DeviceBuffer(n * 16)
            d_b = la.DeviceBuffer(n * 16)
            ctx.extend_dev(d_rays.ptr, n, d_a.ptr)
            ctx.extend_dev_noalloc(d_rays.ptr, n, d_b.ptr)
            a = d_a.download(la.HIT_DTYPE, n)
            b = d_b.download(la.HIT_DTYPE, n)
            assert (a["obj"] != la.LAM_NONE).mean() > 0.05
            assert np.array_equal(_bits(a), _bits(b)), n
            for d in (d_rays, d_a, d_b):
                d.free()
        with pytest.raises(la.LambdaError) as e:
            ctx.extend_dev_noalloc(d_rays.ptr, (1 << 31) + 1, d_a.ptr)
        assert e.value.code == la.LAM_E_LIMIT
        ctx.extend_dev_noalloc(None, 0, None)  # empty batch: nothing to do
    finally:
        ctx.close()
