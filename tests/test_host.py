"""Host side: C ABI exports, error codes without a GPU, tone map, PNG and
the saveImage naming rule."""
import ctypes as C
import math
import struct
import zlib

import numpy as np
import pytest

import lambda_amd as la


def test_library_exports_every_declared_function():
    L = la.lib()
    names = la.declared_functions()
    assert "lam_render" in names and "lamh_load_nls" in names and len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_null_arguments_are_errors_not_crashes():
    L = la.lib()
    assert L.lam_extend(None, None, 0, None) == la.LAM_E_INVALID
    assert L.lam_render(None, None, None, None) == la.LAM_E_INVALID
    assert L.lam_trace_paths(None, None, None, None, 0, None) == la.LAM_E_INVALID
    L.lam_destroy(None)
    assert L.lam_create(None, 0) == la.LAM_E_INVALID


def test_tone_map_closed_form():
    rgb = np.array([[[0.1, 0.2, 0.3], [1.0, 0.5, 0.25]]], np.float32)
    out = la.tone_map(rgb, 0.18)
    Y = rgb @ np.array([0.2126, 0.7152, 0.0722])
    ybar = math.exp(np.mean(np.log(np.maximum(Y, np.finfo(np.float32).eps))))
    x = 10 * rgb * (0.18 / ybar)
    want = (x * (0.3 * x + 0.14)) / (x * (0.3 * x + 2.8) + 10)  # koiFilmic (cg.hpp:181-194)
    np.testing.assert_allclose(out, want, rtol=1e-5)


def test_png_round_trip(tmp_path):
    rgb = np.random.default_rng(0).uniform(0, 1, (7, 5, 3)).astype(np.float32)
    p = tmp_path / "x.png"
    la.write_png(p, rgb)
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h = struct.unpack(">II", data[16:24])
    assert (w, h) == (5, 7)
    idat = data.index(b"IDAT")
    n = struct.unpack(">I", data[idat - 4:idat])[0]
    raw = zlib.decompress(data[idat + 4:idat + 4 + n])
    px = np.frombuffer(raw, np.uint8).reshape(7, 1 + 15)[:, 1:].reshape(7, 5, 3)
    s = np.where(rgb > 0.0031308, 1.055 * np.power(rgb, 1 / 2.4) - 0.055, 12.92 * rgb)
    assert np.abs(px.astype(int) - np.clip((255 * s + 0.5).astype(int), 0, 255)).max() <= 1


def test_save_image_name_rule():
    # Renderer::saveImage (src/renderer.cpp:182-185)
    assert la.image_name("scenes/box.nls", 16, 8, 0.8) == "render/box-16spp-8b-0.80p.png"
    assert la.image_name("scenes/box.nls", 4, 4, 0.5, "-raw") == "render/box-4spp-4b-0.50p-raw.png"


def test_png8_matches_float_png(tmp_path):
    """lamh_write_png8 of the bytes lamh_write_png derives gives the same file"""
    rgb = np.random.default_rng(1).uniform(0, 1, (9, 6, 3)).astype(np.float32)
    a, b = tmp_path / "a.png", tmp_path / "b.png"
    la.write_png(a, rgb)
    data = a.read_bytes()
    idat = data.index(b"IDAT")
    n = struct.unpack(">I", data[idat - 4:idat])[0]
    px = np.frombuffer(zlib.decompress(data[idat + 4:idat + 4 + n]), np.uint8).reshape(9, 1 + 18)[:, 1:]
    la.write_png8(b, px.reshape(9, 6, 3))
    assert b.read_bytes() == data
