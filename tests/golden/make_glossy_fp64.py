"""Writes glossy_interior_64_fp64.npz: the FP64 oracle's render of the
guiding workload's scene (the glossy interior of lambda_amd.scenes with its
seed 4 and its default cabinet layout: 200 Blinn rounded boxes, the light
hidden on a rimmed cabinet top) at 64x64, 4096 samples per
pixel, 32 bounces, p = 0.8, hero 4 — the reference image against which the
guided and plain GPU renders are judged for bias (tests/test_guide.py).

Stored: `mean` (64 x 64 x 3, linear RGB, top row first), `var_of_mean` (the
variance of that mean, from four independent 1024-spp quarters), and the
render settings.  Run from the repository root:
    python tests/golden/make_glossy_fp64.py
(about two minutes on 8 cores)."""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import lambda_amd as la  # noqa: E402
from lambda_amd import scenes  # noqa: E402
from oracle.binding import Oracle  # noqa: E402

SPP, QUARTERS, BOUNCES, P, SEED, WIDTH = 4096, 4, 32, 0.8, 77, 64


def main():
    with tempfile.TemporaryDirectory() as d:
        sc = la.load_nls(scenes.glossy_interior(d, seed=4, width=WIDTH))
        orc = Oracle(sc, f64=True)
        q = SPP // QUARTERS
        parts = [orc.render(la.params(spp=q, max_scatterings=BOUNCES, sample_p=P, hero=sc.hero, seed=SEED,
                                      spp_offset=k * q)).astype(np.float64) for k in range(QUARTERS)]
    parts = np.stack(parts)
    mean = parts.mean(axis=0)
    var_of_mean = parts.var(axis=0, ddof=1) / QUARTERS
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "glossy_interior_64_fp64.npz")
    np.savez_compressed(out, mean=mean.astype(np.float32), var_of_mean=var_of_mean.astype(np.float32),
                        spp=SPP, bounces=BOUNCES, sample_p=P, seed=SEED, width=WIDTH, scene_seed=4)
    print(out, mean.mean(), np.sqrt(var_of_mean).mean())


if __name__ == "__main__":
    main()
