"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (liboracle32.so /
liboracle64.so built by oracle/Makefile).  Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def _load(f64):
    key = bool(f64)
    if key in _LIBS:
        return _LIBS[key]
    path = os.path.join(_HERE, "liboracle64.so" if f64 else "liboracle32.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} not built; run `make -C oracle`")
    L = C.CDLL(path)
    vp, sz, u32p, fp = C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    L.orc_scene_new.restype = vp
    L.orc_scene_new.argtypes = [vp]
    L.orc_scene_free.restype = None
    L.orc_scene_free.argtypes = [vp]
    L.orc_bvh_stats.restype = None
    L.orc_bvh_stats.argtypes = [vp, C.POINTER(C.c_int)]
    L.orc_closest.restype = C.c_int
    L.orc_closest.argtypes = [vp, vp, sz, vp, C.c_int]
    L.orc_occluded.restype = C.c_int
    L.orc_occluded.argtypes = [vp, vp, sz, u32p, C.c_int]
    L.orc_trace_paths.restype = C.c_int
    L.orc_trace_paths.argtypes = [vp, vp, u32p, u32p, sz, fp]
    L.orc_count_rays.restype = C.c_int
    L.orc_count_rays.argtypes = [vp, vp, u32p, u32p, sz, C.POINTER(C.c_uint64)]
    L.orc_render.restype = C.c_int
    L.orc_render.argtypes = [vp, vp, fp, fp, C.c_int, C.c_int, C.c_int]
    i32p, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    L.orc_jvp.restype = C.c_int
    L.orc_jvp.argtypes = [vp, vp, i32p, i32p, i32p, C.c_uint32, C.c_int, dp, fp, C.c_int, dp]
    L.orc_math.restype = None
    L.orc_math.argtypes = [C.c_int, fp, fp, sz, fp]
    L.orc_rng.restype = None
    L.orc_rng.argtypes = [C.c_uint64, u32p, u32p, u32p, sz, u32p]
    _LIBS[key] = L
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Oracle:
    """CPU restatement bound to one flattened scene (a lambda_amd.Scene)."""

    def __init__(self, scene, f64=False):
        self.scene = scene  # keeps the description alive
        self.L = _load(f64)
        self.f64 = f64
        self._h = C.c_void_p(self.L.orc_scene_new(C.cast(scene.desc_ptr, C.c_void_p)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.L.orc_scene_free(self._h)
            self._h = C.c_void_p()

    def bvh_stats(self):
        out = (C.c_int * 4)()
        self.L.orc_bvh_stats(self._h, out)
        return dict(nodes=out[0], interior=out[1], leaves=out[2], depth=out[3])

    def closest(self, rays, brute=False):
        from lambda_amd import HIT_DTYPE, RAY_DTYPE
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        rc = self.L.orc_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data, 0 if brute else 1)
        if rc != 0:
            raise RuntimeError("oracle traversal stack overflow")
        return hits

    def occluded(self, rays, brute=False):
        from lambda_amd import RAY_DTYPE
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros(len(rays), dtype=np.uint32)
        rc = self.L.orc_occluded(self._h, rays.ctypes.data, len(rays), _u(out), 0 if brute else 1)
        if rc != 0:
            raise RuntimeError("oracle traversal stack overflow")
        return out

    def trace_paths(self, p, pixel, sample):
        pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        out = np.zeros((len(pixel), p.hero), dtype=np.float32)
        rc = self.L.orc_trace_paths(self._h, C.cast(C.pointer(p), C.c_void_p), _u(pixel), _u(sample),
                                    len(pixel), _f(out))
        if rc != 0:
            raise RuntimeError("oracle trace_paths failed")
        return out

    def count_rays(self, p, pixel, sample):
        """(extend, shadow) rays tracePath traces for these paths"""
        pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        out = np.zeros(2, dtype=np.uint64)
        rc = self.L.orc_count_rays(self._h, C.cast(C.pointer(p), C.c_void_p), _u(pixel), _u(sample), len(pixel),
                                   out.ctypes.data_as(C.POINTER(C.c_uint64)))
        if rc != 0:
            raise RuntimeError("oracle count_rays failed")
        return int(out[0]), int(out[1])

    def render(self, p, threads=0, row_step=1, row_phase=0, want_xyz=False):
        d = self.scene.desc
        w, h = d.camera.width, d.camera.height
        rgb = np.zeros((h, w, 3), dtype=np.float32)
        xyz = np.zeros((h, w, 3), dtype=np.float32) if want_xyz else None
        rc = self.L.orc_render(self._h, C.cast(C.pointer(p), C.c_void_p), _f(rgb),
                               _f(xyz) if want_xyz else None, int(threads), int(row_step), int(row_phase))
        if rc != 0:
            raise RuntimeError("oracle render failed")
        return (rgb, xyz) if want_xyz else rgb


    def jvp(self, p, dp, v, dloss, threads=0):
        """Directional derivatives of the image loss Σ dloss·rgb along the rows
        of v (K x n_params, K <= 16) by forward-mode tangents; dp is a
        lambda_amd.DiffParams."""
        v = np.ascontiguousarray(np.atleast_2d(v), dtype=np.float64)
        K = v.shape[0]
        if v.shape[1] != dp.n_params or not 1 <= K <= 16:
            raise ValueError(f"tangents have shape {v.shape}")
        g = np.ascontiguousarray(dloss, dtype=np.float32)
        out = np.zeros(K, dtype=np.float64)
        i32 = C.POINTER(C.c_int32)
        rc = self.L.orc_jvp(self._h, C.cast(C.pointer(p), C.c_void_p), dp.mat_d.ctypes.data_as(i32),
                            dp.mat_ior.ctypes.data_as(i32), dp.mat_albedo.ctypes.data_as(i32), dp.n_params, K,
                            v.ctypes.data_as(C.POINTER(C.c_double)), _f(g), int(threads),
                            out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError("oracle jvp failed")
        return out


def math(fn, a, b=None):
    """The oracle's float elementary functions: 0 sin, 1 cos, 2 pow, 3 atan2."""
    L = _load(False)
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(a if b is None else b, dtype=np.float32)
    out = np.zeros_like(a)
    L.orc_math(int(fn), _f(a), _f(b), len(a), _f(out))
    return out


def rng(seed, pixel, sample, dim):
    L = _load(False)
    pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    dim = np.ascontiguousarray(dim, dtype=np.uint32)
    out = np.zeros(len(pixel), dtype=np.uint32)
    L.orc_rng(int(seed), _u(pixel), _u(sample), _u(dim), len(pixel), _u(out))
    return out
