"""The lambda CLI (lambda_amd/csrc/host/main.cpp), the reference's
src/main.cpp:13-47: `lambda <scenes/NAME.nls> [-s SPP -b BOUNCES -p PROB -Y MID]`
renders through the Renderer class over the C ABI and writes
render/NAME-<spp>spp-<b>b-<p>p-raw.png and the tone-mapped
render/NAME-<spp>spp-<b>b-<p>p.png (src/renderer.cpp:179-191).

CPU: argument and load errors (no GPU call is made before the scene loads).
GPU: both PNGs equal, byte for byte, what the Python binding writes for
lam_render with the same parameters and the host tone map.
"""
import os
import subprocess

import numpy as np
import pytest

import lambda_amd as la
from lambda_amd import scenes

CLI = os.path.join(os.path.dirname(la.LIB_PATH), "lambda")


def _run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.fixture
def workdir(tmp_path):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is not built; run make -C lambda_amd")
    (tmp_path / "render").mkdir()
    src = scenes.box_room(str(tmp_path / "scenes"), seed=1, width=48)
    name = os.path.basename(src)
    return tmp_path, os.path.join("scenes", name)


@pytest.mark.parametrize("args", [[], ["-s", "x", "scenes/a.nls"], ["-q", "scenes/a.nls"]])
def test_cli_usage(workdir, args):
    d, _ = workdir
    r = _run(args, d)
    assert r.returncode == 1 and "Usage: ./lambda <fileName> [args]" in r.stdout


def test_cli_missing_scene(workdir):
    d, _ = workdir
    r = _run(["scenes/none.nls", "-s", "2"], d)
    assert r.returncode == 1 and "Failed to load scene." in r.stdout


@pytest.mark.gpu
def test_cli_render_matches_binding(workdir):
    d, rel = workdir
    r = _run([rel, "-s", "4", "-b", "8", "-p", "0.8", "-Y", "0.2"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = d / la.image_name(rel, 4, 8, 0.8, "-raw")
    tm = d / la.image_name(rel, 4, 8, 0.8)
    assert raw.exists() and tm.exists(), os.listdir(d / "render")
    scene = la.load_nls(str(d / rel))
    ctx = la.Context(0)
    try:
        ctx.upload(scene)
        img = ctx.render(la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=scene.hero, seed=0))
    finally:
        ctx.close()
    la.write_png(str(d / "py-raw.png"), img)
    la.write_png(str(d / "py.png"), la.tone_map(img, 0.2))
    assert (d / "py-raw.png").read_bytes() == raw.read_bytes()
    assert (d / "py.png").read_bytes() == tm.read_bytes()
    assert np.asarray(img).mean() > 0
