"""Python binding of the MI355X render loop of lambda (ctypes over the C ABI).

The C ABI is declared in include/lambda_amd.h (device side) and
include/lambda_host.h (loader, tone map, PNG).  This module mirrors the
reference's Renderer interface (include/renderer.hpp:32-52):

    r = Renderer(); r.SPP = 16; r.MAX_SCATTERINGS = 8; r.SAMPLE_P = 0.8
    scene = load_nls("scenes/box.nls")
    raw = r.render(scene)          # Renderer::render  -> linear RGB image
    img = r.toneMap(raw, 0.12)     # Renderer::toneMap
    r.saveImage(raw, "scenes/box.nls", "-raw")

There is no CPU fallback: every call goes to liblambda_amd.so, which must be
built (lambda_amd/Makefile, or __graft_entry__.build()); a missing library
raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblambda_amd.so")
DATA_DIR = os.path.join(_HERE, "data")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

LAM_OK = 0
LAM_E_INVALID = -1
LAM_E_DEVICE = -2
LAM_E_NOSCENE = -3
LAM_E_LIMIT = -4
LAM_E_IO = -5
LAM_E_PARSE = -6
LAM_E_BVH = -7
LAM_NONE = 0xFFFFFFFF

LAM_FLAG_MEGAKERNEL = 0x1
LAM_FLAG_NO_GRAPH = 0x2
LAM_FLAG_PROFILE = 0x4
LAM_FLAG_NO_TAIL = 0x8
LAM_FLAG_NO_PIPE = 0x10
LAM_FLAG_ACCUMULATE = 0x20
LAM_FLAG_PROFILE_FETCH = 0x40
LAM_UPLOAD_GPU_BVH = 0x1

OBJ_SPHERE, OBJ_PLANE, OBJ_TRIMESH = 1, 2, 3
(MAT_NONE, MAT_LAMBERTIAN, MAT_BLINN, MAT_THINFILM, MAT_MICROFACET, MAT_EMITTER,
 MAT_DIREMITTER) = range(7)
LIGHT_AMBIENT, LIGHT_POINT, LIGHT_DIR, LIGHT_SPHERE, LIGHT_QUADDIR = range(5)

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("skip", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("tri", "<u4"), ("obj", "<u4"), ("prim", "<u4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16


class LambdaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


# ---- C structures (layout of include/lambda_amd.h) -------------------------
class Grid(C.Structure):
    _fields_ = [("n", C.c_uint32), ("inf", C.c_float), ("sup", C.c_float)]


class Xform(C.Structure):
    _fields_ = [("M", C.c_float * 12), ("Minv", C.c_float * 12)]


class Object(C.Structure):
    _fields_ = [("type", C.c_uint32), ("material", C.c_uint32), ("mesh", C.c_uint32),
                ("pad", C.c_uint32), ("T", Xform)]


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("spec", C.c_int32 * 5), ("alpha", C.c_float),
                ("d", C.c_float), ("p_d", C.c_float), ("p_spec", C.c_float), ("w_r", C.c_float),
                ("w_t", C.c_float), ("specular", C.c_float), ("pad", C.c_float)]


class Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("object", C.c_uint32), ("spec", C.c_int32),
                ("size", C.c_float), ("inv_area", C.c_float), ("specular", C.c_float),
                ("pos", C.c_float * 3), ("dir", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("x", C.c_float * 3), ("y", C.c_float * 3),
                ("z", C.c_float * 3), ("D", C.c_float), ("delta", C.c_float), ("w", C.c_float),
                ("h", C.c_float), ("dof", C.c_float), ("fov", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("grid", Grid), ("n_spectra", C.c_uint32), ("spectra", C.POINTER(C.c_float)),
                ("cmf", C.POINTER(C.c_float)), ("n_objects", C.c_uint32),
                ("objects", C.POINTER(Object)), ("n_materials", C.c_uint32),
                ("materials", C.POINTER(Material)), ("n_lights", C.c_uint32),
                ("lights", C.POINTER(Light)), ("n_tris", C.c_uint32),
                ("tri_v", C.POINTER(C.c_float)), ("tri_n", C.POINTER(C.c_float)),
                ("tri_gn", C.POINTER(C.c_float)), ("tri_obj", C.POINTER(C.c_uint32)),
                ("tri_prim", C.POINTER(C.c_uint32)), ("camera", Camera)]


class DiffParamsC(C.Structure):
    _fields_ = [("n_params", C.c_uint32), ("n_materials", C.c_uint32), ("mat_d", C.POINTER(C.c_int32)),
                ("mat_ior", C.POINTER(C.c_int32)), ("mat_albedo", C.POINTER(C.c_int32))]


class GuideParams(C.Structure):
    _fields_ = [("alpha", C.c_float), ("train_passes", C.c_uint32), ("steps_per_pass", C.c_uint32),
                ("batch", C.c_uint32), ("lr", C.c_float), ("seed", C.c_uint64), ("materials", C.c_uint32),
                ("max_vertex", C.c_uint32), ("lobes", C.c_uint32)]


GUIDE_LAMBERTIAN, GUIDE_BLINN_KD, GUIDE_BLINN_SPEC = 0x1, 0x2, 0x4


def guide_params(alpha=0.6, train_passes=32, steps_per_pass=128, batch=16384, lr=5e-3, seed=1, materials=0,
                 max_vertex=0, lobes=0):
    """lam_guide_params: guide-branch probability, online-training schedule,
    guided material types (bit 1 << MAT_*; 0 = every supported type),
    guided vertices (bounce index < max_vertex; 0 = every vertex), guided
    lobes (GUIDE_* bits; 0 = the two diffuse lobes, Lambertian and Blinn Kd)."""
    return GuideParams(float(alpha), int(train_passes), int(steps_per_pass), int(batch), float(lr), int(seed),
                       int(materials), int(max_vertex), int(lobes))


GUIDE_IN, GUIDE_OUT, GUIDE_LOBES = 6, 40, 8


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_scatterings", C.c_uint32), ("sample_p", C.c_float),
                ("hero", C.c_uint32), ("seed", C.c_uint64), ("spp_offset", C.c_uint32),
                ("flags", C.c_uint32)]


class SceneStats(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32), ("n_tris", C.c_uint32),
                ("depth", C.c_uint32), ("stack_bound", C.c_uint32), ("n_analytic", C.c_uint32),
                ("device_bytes", C.c_uint64), ("sah_cost", C.c_double)]


class WaveConfig(C.Structure):
    """lam_wave_config"""
    _fields_ = [("lanes", C.c_uint32), ("tail_threshold", C.c_uint32), ("tail_div", C.c_uint32),
                ("max_iters", C.c_uint32), ("trace_per_cu", C.c_uint32), ("batch_per_cu", C.c_uint32),
                ("finish_per_cu", C.c_uint32), ("shade_per_cu", C.c_uint32), ("samples_per_pass", C.c_uint32)]


class Timing(C.Structure):
    _fields_ = [("extend_ms", C.c_double), ("shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("raygen_ms", C.c_double), ("film_ms", C.c_double), ("other_ms", C.c_double),
                ("extend_launches", C.c_uint64), ("shadow_launches", C.c_uint64),
                ("shade_launches", C.c_uint64), ("extend_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("render_ms", C.c_double),
                ("guide_records", C.c_uint64), ("guide_train_ms", C.c_double),
                ("guide_dropped", C.c_uint64), ("guide_zero_pdf", C.c_uint64),
                ("guide_loss_first", C.c_double), ("guide_loss_last", C.c_double),
                ("extend_nodes", C.c_uint64), ("extend_tris", C.c_uint64), ("extend_traced", C.c_uint64),
                ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64), ("shadow_traced", C.c_uint64),
                ("pass_samples", C.c_uint32), ("passes", C.c_uint32)]


def params(spp=4, max_scatterings=4, sample_p=0.5, hero=16, seed=0, spp_offset=0, flags=0):
    """lam_render_params with the reference's Renderer defaults
    (include/renderer.hpp:35-37)."""
    return RenderParams(int(spp), int(max_scatterings), float(sample_p), int(hero), int(seed),
                        int(spp_offset), int(flags))


_LIB = None


def lib():
    """The loaded liblambda_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C lambda_amd` "
                          "or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, sz, u32p, fp = C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    sig = {
        "lam_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
        "lam_destroy": (None, [vp]),
        "lam_last_error": (C.c_char_p, [vp]),
        "lam_version": (C.c_char_p, []),
        "lam_scene_upload": (C.c_int, [vp, vp]),
        "lam_get_scene_stats": (C.c_int, [vp, C.POINTER(SceneStats)]),
        "lam_render": (C.c_int, [vp, C.POINTER(RenderParams), fp, fp]),
        "lam_trace_paths": (C.c_int, [vp, C.POINTER(RenderParams), u32p, u32p, sz, fp]),
        "lam_adjoint": (C.c_int, [vp, C.POINTER(RenderParams), C.POINTER(DiffParamsC), fp, fp]),
        "lam_render_guided": (C.c_int, [vp, C.POINTER(RenderParams), C.POINTER(GuideParams), fp, fp]),
        "lam_guide_param_count": (C.c_size_t, []),
        "lam_guide_reset": (C.c_int, [vp, C.c_uint64]),
        "lam_guide_set_params": (C.c_int, [vp, fp]),
        "lam_guide_get_params": (C.c_int, [vp, fp]),
        "lam_guide_infer": (C.c_int, [vp, fp, sz, fp]),
        "lam_guide_loss_grad": (C.c_int, [vp, fp, sz, fp, fp, fp]),
        "lam_guide_train": (C.c_int, [vp, fp, sz, C.c_uint32, C.c_float, fp]),
        "lam_extend": (C.c_int, [vp, vp, sz, vp]),
        "lam_occluded": (C.c_int, [vp, vp, sz, u32p]),
        "lam_extend_dev": (C.c_int, [vp, vp, sz, vp, vp]),
        "lam_extend_dev_timed": (C.c_int, [vp, vp, sz, vp, C.c_uint32, C.POINTER(C.c_double)]),
        "lam_extend_dev_noalloc": (C.c_int, [vp, vp, sz, vp]),
        "lam_occluded_dev": (C.c_int, [vp, vp, sz, vp, vp]),
        "lam_occluded_dev_timed": (C.c_int, [vp, vp, sz, vp, C.c_uint32, C.POINTER(C.c_double)]),
        "lam_get_timing": (C.c_int, [vp, C.POINTER(Timing)]),
        "lam_reset_timing": (C.c_int, [vp]),
        "lam_get_wave_config": (C.c_int, [vp, C.POINTER(WaveConfig)]),
        "lam_set_camera": (C.c_int, [vp, C.POINTER(Camera)]),
        "lam_scene_upload_ex": (C.c_int, [vp, vp, C.c_uint32]),
        "lam_film_get": (C.c_int, [vp, fp, u32p, u32p]),
        "lam_film_set": (C.c_int, [vp, fp, C.c_uint32, C.c_uint32]),
        "lam_tone_map": (C.c_int, [vp, C.c_float, C.POINTER(C.c_uint8)]),
        "lam_set_wave_config": (C.c_int, [vp, C.POINTER(WaveConfig)]),
        "lam_debug_math": (C.c_int, [vp, C.c_int, fp, fp, sz, fp]),
        "lam_debug_radix_sort": (C.c_int, [vp, u32p, u32p, sz, C.c_uint32, u32p, u32p]),
        "lamh_load_nls": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, sz]),
        "lamh_load_nls_text": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, sz]),
        "lamh_scene_free": (None, [vp]),
        "lamh_scene_desc": (C.POINTER(SceneDesc), [vp]),
        "lamh_scene_hero": (C.c_uint32, [vp]),
        "lamh_object_name": (C.c_char_p, [vp, C.c_uint32]),
        "lamh_material_name": (C.c_char_p, [vp, C.c_uint32]),
        "lamh_tone_map": (C.c_int, [fp, C.c_uint32, C.c_uint32, C.c_float, fp]),
        "lamh_write_png": (C.c_int, [C.c_char_p, fp, C.c_uint32, C.c_uint32]),
        "lamh_write_png8": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32]),
        "lamh_image_name": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_float, C.c_char_p,
                                      C.c_char_p, sz]),
        "lamh_bvh8_check": (C.c_int, [fp, C.c_uint32, C.POINTER(SceneStats), C.POINTER(C.c_int)]),
        "lamh_spectrum_grad_to_samples": (C.c_int, [C.POINTER(Grid), fp, sz, C.c_float, fp, fp]),
        "lamh_trace_counts": (C.c_int, [vp, vp, sz, C.c_int, vp, u32p, u32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def declared_functions():
    """Names of every function declared in include/*.h (for the export test)."""
    names = []
    for h in sorted(os.listdir(INCLUDE_DIR)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE_DIR, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(lamh?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _carr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


_HIP = None


def _hip():
    """libamdhip64 for the device buffers of the `_dev` entry points"""
    global _HIP
    if _HIP is None:
        lib()  # its NEEDED libamdhip64.so.7 is the runtime instance to share
        # (another name could load a second copy of the runtime, e.g. beside
        # torch's own libamdhip64, which then reports no device)
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                h = C.CDLL(name)
            except OSError:
                continue
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMalloc.restype = C.c_int
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemcpy.restype = C.c_int
            h.hipFree.argtypes = [C.c_void_p]
            h.hipFree.restype = C.c_int
            h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
            h.hipHostMalloc.restype = C.c_int
            h.hipHostFree.argtypes = [C.c_void_p]
            h.hipHostFree.restype = C.c_int
            _HIP = h
            break
        else:
            raise RuntimeError("libamdhip64.so not found")
    return _HIP


class _HostBuffer:
    """A page-locked host buffer seen by numpy (np.asarray(buf) has it as
    its base); it goes back to its pool when the last array viewing it
    is collected."""

    def __init__(self, pool, ptr, nbytes, shape, dtype):
        self._pool, self._ptr, self._nbytes = pool, ptr, nbytes
        self.__array_interface__ = {"shape": tuple(shape), "typestr": np.dtype(dtype).str,
                                    "data": (ptr, False), "version": 3}

    def __del__(self):
        self._pool.release(self._ptr, self._nbytes)


class _PinnedPool:
    """Page-locked (hipHostMalloc) host memory for the images the device
    returns: the device-to-host copy then runs as one DMA into the caller's
    array, without staging through pageable memory or first-touch page
    faults (about 1.7 ms of a 1080p frame's 25 MB read-back).  Freed
    buffers are kept for the next request of the same size, at most
    `keep` per size."""

    def __init__(self, keep=4):
        self._free = {}
        self._keep = keep

    def array(self, shape, dtype=np.float32):
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if nbytes == 0:
            return np.zeros(shape, dtype)
        free = self._free.get(nbytes)
        if free:
            ptr = free.pop()
        else:
            ptr = C.c_void_p()
            if _hip().hipHostMalloc(C.byref(ptr), nbytes, 0) != 0 or not ptr.value:
                return np.empty(shape, dtype)  # ordinary memory when none is left
            ptr = ptr.value
        return np.asarray(_HostBuffer(self, ptr, nbytes, shape, dtype))

    def release(self, ptr, nbytes):
        free = self._free.setdefault(nbytes, [])
        if len(free) < self._keep:
            free.append(ptr)
        elif _HIP is not None:
            _HIP.hipHostFree(C.c_void_p(ptr))


_PINNED = _PinnedPool()


class DeviceBuffer:
    """A block of device memory (hipMalloc) for the `_dev` entry points, on
    the current HIP device."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        rc = _hip().hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 16)))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {rc}")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if _hip().hipMemcpy(self.ptr, a.ctypes.data, C.c_size_t(a.nbytes), 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to the device failed")

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read past the buffer")
        if _hip().hipMemcpy(out.ctypes.data, self.ptr, C.c_size_t(out.nbytes), 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def free(self):
        if self.ptr.value:
            _hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """A loaded .nls scene (nl::cg::scene) and its flattened description."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self._lib = lib()

    @classmethod
    def load(cls, path, data_dir=None):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.lamh_load_nls(str(path).encode(), (data_dir or DATA_DIR).encode(), C.byref(h), err, 1024)
        if rc != LAM_OK:
            raise LambdaError(rc, err.value.decode())
        return cls(h.value)

    @classmethod
    def from_text(cls, text, data_dir=None):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.lamh_load_nls_text(text.encode(), (data_dir or DATA_DIR).encode(), C.byref(h), err, 1024)
        if rc != LAM_OK:
            raise LambdaError(rc, err.value.decode())
        return cls(h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lamh_scene_free(h)
            self._h = C.c_void_p()

    @property
    def handle(self):
        return self._h

    @property
    def desc_ptr(self):
        return self._lib.lamh_scene_desc(self._h)

    @property
    def desc(self):
        return self.desc_ptr.contents

    @property
    def hero(self):
        return int(self._lib.lamh_scene_hero(self._h))

    @property
    def width(self):
        return int(self.desc.camera.width)

    @property
    def height(self):
        return int(self.desc.camera.height)

    @property
    def n_tris(self):
        return int(self.desc.n_tris)

    def tri_v(self):
        d = self.desc
        return _carr(d.tri_v, 9 * d.n_tris, np.float32).reshape(-1, 3, 3)

    def tri_obj(self):
        d = self.desc
        return _carr(d.tri_obj, d.n_tris, np.uint32)

    def tri_prim(self):
        d = self.desc
        return _carr(d.tri_prim, d.n_tris, np.uint32)

    def objects(self):
        d = self.desc
        return [d.objects[i] for i in range(d.n_objects)]

    def materials(self):
        d = self.desc
        return [d.materials[i] for i in range(d.n_materials)]

    def lights(self):
        d = self.desc
        return [d.lights[i] for i in range(d.n_lights)]

    def spectra(self):
        d = self.desc
        return _carr(d.spectra, d.n_spectra * d.grid.n, np.float32).reshape(d.n_spectra, d.grid.n)

    def cmf(self):
        d = self.desc
        return _carr(d.cmf, 3 * d.grid.n, np.float32).reshape(3, d.grid.n)

    def object_name(self, i):
        v = self._lib.lamh_object_name(self._h, i)
        return None if v is None else v.decode()

    def material_name(self, i):
        v = self._lib.lamh_material_name(self._h, i)
        return None if v is None else v.decode()


class SceneVariant:
    """A scene description whose materials and spectrum table replace those of
    a base scene (geometry, lights and camera are the base scene's, which
    must outlive the variant).  Accepted wherever a Scene is uploaded."""

    def __init__(self, base, materials, spectra):
        self.base = base
        self._mats = (Material * len(materials))(*materials)
        self._spec = np.ascontiguousarray(spectra, dtype=np.float32)
        d = SceneDesc()
        C.memmove(C.byref(d), base.desc_ptr, C.sizeof(SceneDesc))
        d.n_spectra = self._spec.shape[0]
        d.spectra = _fptr(self._spec)
        d.n_materials = len(materials)
        d.materials = C.cast(self._mats, C.POINTER(Material))
        self._desc = d

    @property
    def desc_ptr(self):
        return C.pointer(self._desc)

    @property
    def desc(self):
        return self._desc

    @property
    def hero(self):
        return self.base.hero

    @property
    def width(self):
        return int(self._desc.camera.width)

    @property
    def height(self):
        return int(self._desc.camera.height)

    def materials(self):
        return [self._mats[i] for i in range(len(self._mats))]

    def spectra(self):
        return self._spec


def _copy_material(m):
    c = Material()
    C.memmove(C.byref(c), C.byref(m), C.sizeof(Material))
    return c


class DiffParams:
    """The differentiated parameters (lam_diff_params) of one scene: a slot
    for the thin-film thickness d of each material in `d`, grid.n slots for
    the film-IOR coefficients of each material in `ior` (thin film) and for
    the albedo coefficients of each material in `albedo` (Lambertian), in
    that order.  Materials are given by index or by name."""

    def __init__(self, scene, d=(), ior=(), albedo=()):
        desc = scene.desc
        nm, gn = desc.n_materials, desc.grid.n
        mats = [desc.materials[i] for i in range(nm)]
        names = {}
        base = scene.base if isinstance(scene, SceneVariant) else scene
        for i in range(nm):
            n = base.material_name(i)
            if n is not None:
                names[n] = i

        def idx(m):
            if isinstance(m, str):
                if m not in names:
                    raise KeyError(f"no material named {m!r}")
                return names[m]
            m = int(m)
            if not 0 <= m < nm:
                raise IndexError(f"material {m} out of range")
            return m

        self.n_materials, self.grid_n = nm, gn
        self.mat_d = np.full(nm, -1, dtype=np.int32)
        self.mat_ior = np.full(nm, -1, dtype=np.int32)
        self.mat_albedo = np.full(nm, -1, dtype=np.int32)
        off = 0
        for lst, arr, kind, width in ((d, self.mat_d, MAT_THINFILM, 1), (ior, self.mat_ior, MAT_THINFILM, gn),
                                      (albedo, self.mat_albedo, MAT_LAMBERTIAN, gn)):
            for m in lst:
                i = idx(m)
                if mats[i].type != kind:
                    raise ValueError(f"material {i} has type {mats[i].type}, not {kind}")
                if arr[i] >= 0:
                    raise ValueError(f"material {i} listed twice")
                arr[i] = off
                off += width
        self.n_params = off

    def c_struct(self):
        i32 = C.POINTER(C.c_int32)
        return DiffParamsC(self.n_params, len(self.mat_d), self.mat_d.ctypes.data_as(i32),
                           self.mat_ior.ctypes.data_as(i32), self.mat_albedo.ctypes.data_as(i32))

    def values(self, scene):
        """The parameter vector θ of a scene (or variant)."""
        mats = scene.materials()
        spec = scene.spectra()
        th = np.zeros(self.n_params, dtype=np.float32)
        gn = self.grid_n
        for m in range(self.n_materials):
            if self.mat_d[m] >= 0:
                th[self.mat_d[m]] = mats[m].d
            if self.mat_ior[m] >= 0:
                th[self.mat_ior[m]:self.mat_ior[m] + gn] = spec[mats[m].spec[0]]
            if self.mat_albedo[m] >= 0:
                th[self.mat_albedo[m]:self.mat_albedo[m] + gn] = spec[mats[m].spec[0]]
        return th

    def apply(self, scene, theta):
        """A SceneVariant of `scene` with the parameters set to θ; each
        differentiated spectrum gets a row of its own."""
        base = scene.base if isinstance(scene, SceneVariant) else scene
        theta = np.asarray(theta, dtype=np.float32)
        if theta.shape != (self.n_params,):
            raise ValueError(f"theta has shape {theta.shape}, expected ({self.n_params},)")
        mats = [_copy_material(m) for m in scene.materials()]
        spec = [np.asarray(scene.spectra(), dtype=np.float32)]
        rows = spec[0].shape[0]
        gn = self.grid_n
        for m in range(self.n_materials):
            if self.mat_d[m] >= 0:
                mats[m].d = float(theta[self.mat_d[m]])
            for arr in (self.mat_ior, self.mat_albedo):
                if arr[m] >= 0:
                    spec.append(theta[arr[m]:arr[m] + gn][None, :])
                    mats[m].spec[0] = rows
                    rows += 1
        return SceneVariant(base, mats, np.concatenate(spec, axis=0))


def load_nls(path, data_dir=None):
    """nl::cg::load::loadNLS (lib/numli/cg.hpp:2455-2470)."""
    return Scene.load(path, data_dir)


class Context:
    """One lam_ctx: a HIP stream and the HBM copy of one scene."""

    def __init__(self, device=0):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.lam_create(C.byref(h), int(device))
        if rc != LAM_OK:
            raise LambdaError(rc, "lam_create failed (no GPU or HIP runtime error)")
        self._h = h
        self.scene = None
        self._size = (0, 0)  # image size of the current camera

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lam_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        if rc != LAM_OK:
            raise LambdaError(rc, self._lib.lam_last_error(self._h).decode())

    def upload(self, scene, gpu_bvh=False):
        """lam_scene_upload(_ex): gpu_bvh builds the BVH8 on the device"""
        self._check(self._lib.lam_scene_upload_ex(self._h, C.cast(scene.desc_ptr, C.c_void_p),
                                                  LAM_UPLOAD_GPU_BVH if gpu_bvh else 0))
        self.scene = scene
        self._size = (scene.width, scene.height)

    def stats(self):
        s = SceneStats()
        self._check(self._lib.lam_get_scene_stats(self._h, C.byref(s)))
        return s

    def extend(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        self._check(self._lib.lam_extend(self._h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros(len(rays), dtype=np.uint32)
        self._check(self._lib.lam_occluded(self._h, rays.ctypes.data, len(rays), _u32ptr(out)))
        return out

    def extend_dev(self, rays_ptr, n, hits_ptr, counts_ptr=None):
        self._check(self._lib.lam_extend_dev(self._h, rays_ptr, n, hits_ptr, counts_ptr))

    def occluded_dev(self, rays_ptr, n, occl_ptr, counts_ptr=None):
        self._check(self._lib.lam_occluded_dev(self._h, rays_ptr, n, occl_ptr, counts_ptr))

    def occluded_dev_timed(self, rays_ptr, n, occl_ptr, reps):
        ms = C.c_double()
        self._check(self._lib.lam_occluded_dev_timed(self._h, rays_ptr, n, occl_ptr, int(reps), C.byref(ms)))
        return ms.value

    def extend_dev_noalloc(self, rays_ptr, n, hits_ptr):
        """lam_extend_dev without device allocation (static device scratch)"""
        self._check(self._lib.lam_extend_dev_noalloc(self._h, rays_ptr, n, hits_ptr))

    def extend_dev_timed(self, rays_ptr, n, hits_ptr, reps):
        ms = C.c_double()
        self._check(self._lib.lam_extend_dev_timed(self._h, rays_ptr, n, hits_ptr, int(reps), C.byref(ms)))
        return ms.value

    def trace_paths(self, p, pixel, sample):
        pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        if pixel.shape != sample.shape or pixel.ndim != 1:
            raise ValueError(f"pixel {pixel.shape} and sample {sample.shape} must be 1-D of one length")
        out = np.zeros((len(pixel), p.hero), dtype=np.float32)
        self._check(self._lib.lam_trace_paths(self._h, C.byref(p), _u32ptr(pixel), _u32ptr(sample),
                                              len(pixel), _fptr(out)))
        return out

    def render(self, p, want_xyz=False):
        w, h = self._size
        rgb = _PINNED.array((h, w, 3))
        xyz = _PINNED.array((h, w, 3)) if want_xyz else None
        self._check(self._lib.lam_render(self._h, C.byref(p), _fptr(rgb),
                                         _fptr(xyz) if want_xyz else None))
        return (rgb, xyz) if want_xyz else rgb

    def adjoint(self, p, dp, dloss_drgb):
        """lam_adjoint: gradient of an image loss with respect to the
        parameters of `dp` (a DiffParams) at the uploaded scene, given
        ∂loss/∂rgb (h x w x 3, as render's output)."""
        w, h = self._size
        g = np.ascontiguousarray(dloss_drgb, dtype=np.float32)
        if g.shape != (h, w, 3):
            raise ValueError(f"dloss_drgb has shape {g.shape}")
        if self.scene is not None and dp.n_materials != self.scene.desc.n_materials:
            raise ValueError(f"DiffParams built for {dp.n_materials} materials, the uploaded scene has "
                             f"{self.scene.desc.n_materials}")
        grad = np.zeros(dp.n_params, dtype=np.float32)
        cs = dp.c_struct()
        self._check(self._lib.lam_adjoint(self._h, C.byref(p), C.byref(cs), _fptr(g), _fptr(grad)))
        return grad

    # ---- learned path guiding ----
    def render_guided(self, p, g, want_xyz=False):
        """lam_render_guided: render with the learned proposal (g: GuideParams)."""
        w, h = self._size
        rgb = _PINNED.array((h, w, 3))
        xyz = _PINNED.array((h, w, 3)) if want_xyz else None
        self._check(self._lib.lam_render_guided(self._h, C.byref(p), C.byref(g), _fptr(rgb),
                                                _fptr(xyz) if want_xyz else None))
        return (rgb, xyz) if want_xyz else rgb

    def guide_reset(self, seed=1):
        self._check(self._lib.lam_guide_reset(self._h, int(seed)))

    def guide_params_count(self):
        return int(self._lib.lam_guide_param_count())

    def guide_get_params(self):
        th = np.zeros(self.guide_params_count(), dtype=np.float32)
        self._check(self._lib.lam_guide_get_params(self._h, _fptr(th)))
        return th

    def guide_set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.shape != (self.guide_params_count(),):
            raise ValueError(f"theta has shape {th.shape}")
        self._check(self._lib.lam_guide_set_params(self._h, _fptr(th)))

    def guide_infer(self, inputs):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, GUIDE_IN)
        raw = np.zeros((len(x), GUIDE_OUT), dtype=np.float32)
        self._check(self._lib.lam_guide_infer(self._h, _fptr(x), len(x), _fptr(raw)))
        return raw

    def guide_loss_grad(self, rows, want_raw=False):
        r = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, 10)
        loss = C.c_float()
        grad = np.zeros(self.guide_params_count(), dtype=np.float32)
        raw = np.zeros((len(r), GUIDE_OUT), dtype=np.float32) if want_raw else None
        self._check(self._lib.lam_guide_loss_grad(self._h, _fptr(r), len(r), C.byref(loss), _fptr(grad),
                                                  _fptr(raw) if want_raw else None))
        return (loss.value, grad, raw) if want_raw else (loss.value, grad)

    def guide_train(self, rows, steps, lr):
        r = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, 10)
        loss = C.c_float()
        self._check(self._lib.lam_guide_train(self._h, _fptr(r), len(r), int(steps), float(lr), C.byref(loss)))
        return loss.value

    def timing(self):
        t = Timing()
        self._check(self._lib.lam_get_timing(self._h, C.byref(t)))
        return t

    def reset_timing(self):
        self._check(self._lib.lam_reset_timing(self._h))

    # ---- progressive rendering, checkpoints, device tone map ----
    def film_get(self):
        """checkpoint of the film: (h x w x 8 float32 film, spp, hero)"""
        w, h = self._size
        film = np.zeros((h, w, 8), dtype=np.float32)
        spp, hero = C.c_uint32(), C.c_uint32()
        self._check(self._lib.lam_film_get(self._h, _fptr(film), C.byref(spp), C.byref(hero)))
        return film, spp.value, hero.value

    def film_set(self, film, spp, hero):
        """restore a checkpoint; continue it with LAM_FLAG_ACCUMULATE"""
        w, h = self._size
        film = np.ascontiguousarray(film, dtype=np.float32)
        if film.shape != (h, w, 8):
            raise ValueError(f"film has shape {film.shape}, the camera needs {(h, w, 8)}")
        self._check(self._lib.lam_film_set(self._h, _fptr(film), int(spp), int(hero)))

    def tone_map8(self, y_mid=0.12):
        """lam_tone_map: tone-mapped 8-bit sRGB (h x w x 3 uint8) of the last render"""
        w, h = self._size
        out = np.zeros((h, w, 3), dtype=np.uint8)
        self._check(self._lib.lam_tone_map(self._h, float(y_mid), out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def render_progressive(self, p, chunk, guide=None):
        """Render p.spp samples per pixel in pieces of `chunk` passes (from
        p.spp_offset), yielding (image, samples per pixel so far, estimated
        seconds left) after each piece; the last image equals a one-shot
        render of p bit for bit."""
        import time
        done, t0 = 0, time.perf_counter()
        while done < p.spp:
            n = min(chunk, p.spp - done)
            q = RenderParams(spp=n, max_scatterings=p.max_scatterings, sample_p=p.sample_p, hero=p.hero,
                             seed=p.seed, spp_offset=p.spp_offset + done,
                             flags=p.flags | (LAM_FLAG_ACCUMULATE if done else 0))
            img = self.render(q) if guide is None else self.render_guided(q, guide)
            done += n
            el = time.perf_counter() - t0
            yield img, done, el / done * (p.spp - done)

    def set_camera(self, scene):
        """use `scene`'s camera (a view of the uploaded geometry, e.g. the same
        scene loaded at another resolution) without uploading again"""
        cam = Camera()
        C.memmove(C.byref(cam), C.byref(scene.desc.camera), C.sizeof(Camera))
        self._check(self._lib.lam_set_camera(self._h, C.byref(cam)))
        self._size = (scene.width, scene.height)

    def wave_config(self):
        """the wavefront launch configuration as a dict (lam_get_wave_config)"""
        c = WaveConfig()
        self._check(self._lib.lam_get_wave_config(self._h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in WaveConfig._fields_}

    def set_wave_config(self, **kw):
        """change fields of the launch configuration (lanes, tail_threshold,
        tail_div, max_iters, trace_per_cu, batch_per_cu); images and hits do
        not depend on them"""
        c = WaveConfig(**{**self.wave_config(), **kw})
        self._check(self._lib.lam_set_wave_config(self._h, C.byref(c)))

    def debug_math(self, fn, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        b = np.ascontiguousarray(a if b is None else b, dtype=np.float32)
        out = np.zeros_like(a)
        self._check(self._lib.lam_debug_math(self._h, int(fn), _fptr(a), _fptr(b), len(a), _fptr(out)))
        return out

    def debug_radix_sort(self, keys, vals, bits=32):
        """the device BVH builder's stable radix sort of (keys, vals) by the
        low `bits` bits of the keys (test hook)"""
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        v = np.ascontiguousarray(vals, dtype=np.uint32)
        ko, vo = np.zeros_like(k), np.zeros_like(v)
        self._check(self._lib.lam_debug_radix_sort(self._h, _u32ptr(k), _u32ptr(v), len(k), int(bits),
                                                   _u32ptr(ko), _u32ptr(vo)))
        return ko, vo


def _u32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def host_trace(scene, rays, any_hit=False):
    """lamh_trace_counts: the kernels' BVH8 traversal on the host.  Returns
    (hits or occluded flags, counts n x 2: nodes fetched, triangles tested)."""
    L = lib()
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = len(rays)
    counts = np.zeros((n, 2), np.uint32)
    if any_hit:
        out = np.zeros(n, np.uint32)
        rc = L.lamh_trace_counts(C.cast(scene.desc_ptr, C.c_void_p), rays.ctypes.data, n, 1, None, _u32ptr(out),
                                 _u32ptr(counts))
    else:
        out = np.zeros(n, dtype=HIT_DTYPE)
        rc = L.lamh_trace_counts(C.cast(scene.desc_ptr, C.c_void_p), rays.ctypes.data, n, 0, out.ctypes.data, None,
                                 _u32ptr(counts))
    if rc != LAM_OK:
        raise LambdaError(rc, "lamh_trace_counts failed")
    return out, counts


def tone_map(rgb, y_mid=0.12):
    """Renderer::toneMap (src/renderer.cpp:149-176)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros_like(rgb)
    h, w = rgb.shape[:2]
    rc = lib().lamh_tone_map(_fptr(rgb), w, h, float(y_mid), _fptr(out))
    if rc != LAM_OK:
        raise LambdaError(rc, "tone map failed")
    return out


def spectrum_grad_to_samples(grid, sample_l, dcoef, mult=1.0):
    """Gradient with respect to a spectrum CSV's value column, given the
    gradient with respect to its grid coefficients (the transpose of the
    loader's box filter, lib/numli/cg.hpp:252-272; `mult` is the spectrum's
    "mult").  `grid` is a Grid (Scene.desc.grid); `dcoef` a slot range of
    Context.adjoint's gradient, or rows of them (..., grid.n)."""
    sample_l = np.ascontiguousarray(sample_l, dtype=np.float32)
    dcoef = np.ascontiguousarray(dcoef, dtype=np.float32)
    if dcoef.shape[-1] != grid.n:
        raise ValueError(f"dcoef has {dcoef.shape[-1]} values per row, the grid {grid.n}")
    rows = dcoef.reshape(-1, grid.n)
    out = np.zeros((rows.shape[0], len(sample_l)), dtype=np.float32)
    for r in range(rows.shape[0]):
        rc = lib().lamh_spectrum_grad_to_samples(C.byref(grid), _fptr(sample_l), len(sample_l), float(mult),
                                                 _fptr(rows[r]), _fptr(out[r]))
        if rc != LAM_OK:
            raise LambdaError(rc, "spectrum gradient mapping failed")
    return out.reshape(dcoef.shape[:-1] + (len(sample_l),))


def write_png8(path, rgb8):
    """PNG from 8-bit sRGB (h x w x 3 uint8, e.g. Context.tone_map8's output)."""
    rgb8 = np.ascontiguousarray(rgb8, dtype=np.uint8)
    h, w = rgb8.shape[:2]
    rc = lib().lamh_write_png8(str(path).encode(), rgb8.ctypes.data_as(C.POINTER(C.c_uint8)), w, h)
    if rc != LAM_OK:
        raise LambdaError(rc, f"could not write {path}")


def write_png(path, rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    rc = lib().lamh_write_png(str(path).encode(), _fptr(rgb), w, h)
    if rc != LAM_OK:
        raise LambdaError(rc, f"could not write {path}")


def image_name(scene_path, spp, bounces, p, suffix=""):
    out = C.create_string_buffer(4096)
    rc = lib().lamh_image_name(str(scene_path).encode(), spp, bounces, float(p), suffix.encode(), out, 4096)
    if rc != LAM_OK:
        raise LambdaError(rc, f"no image name for {scene_path}")
    return out.value.decode()


class Renderer:
    """The reference's Renderer (include/renderer.hpp:32-52) over the GPU.

    tracePath is keyed by (pixel, sample) instead of (λ, ray, RNG&): every
    path draws its random numbers from a counter-based RNG addressed by
    (seed, pixel, sample, dimension)."""

    def __init__(self, device=0):
        self.MAX_SCATTERINGS = 4
        self.SPP = 4
        self.SAMPLE_P = 0.5
        self.SEED = 0
        self.FLAGS = 0
        self._ctx = Context(device)
        self._scene = None

    def _params(self, scene):
        return params(self.SPP, self.MAX_SCATTERINGS, self.SAMPLE_P, scene.hero, self.SEED, 0, self.FLAGS)

    def _upload(self, scene):
        if self._scene is not scene:
            self._ctx.upload(scene)
            self._scene = scene

    def render(self, scene):
        self._upload(scene)
        return self._ctx.render(self._params(scene))

    def toneMap(self, buffer, Y_MID=0.12):
        return tone_map(buffer, Y_MID)

    def tracePath(self, scene, pixel, sample):
        self._upload(scene)
        return self._ctx.trace_paths(self._params(scene), [pixel], [sample])[0]

    def saveImage(self, buffer, fpath, suffix=""):
        write_png(image_name(fpath, self.SPP, self.MAX_SCATTERINGS, self.SAMPLE_P, suffix), buffer)
