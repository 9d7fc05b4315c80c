"""gisspa_amd -- MI355X-native GisSPA template-matching picker.

Python view of the C ABI in include/gisspa.h (libgisspa.so, HIP kernels for
gfx950).  The package is plumbing for tests and benchmarks; the product is
the shared library and the ./main CLI built next to it.  Loading fails loudly
if the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgisspa.so")
MAIN_PATH = os.path.join(_HERE, "main")

GS_OK = 0
GS_ERR_CAPACITY = -6
GS_ERR_IO = -7


class GsParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("apix", "phi_step", "kk", "energy", "cs", "highres", "lowres", "diameter",
                 "threshold")] + \
               [(n, ctypes.c_int) for n in
                ("window_size", "phase_flip", "overlap", "norm_type", "invert")] + \
               [(n, ctypes.c_float) for n in
                ("ampconst", "a", "b", "b2", "bfactor", "bfactor2", "bfactor3")]


class GsCtf(ctypes.Structure):
    _fields_ = [("defocus", ctypes.c_float), ("dfdiff", ctypes.c_float), ("dfang", ctypes.c_float)]


class GsParticle(ctypes.Structure):
    _fields_ = [("psi_index", ctypes.c_int), ("view", ctypes.c_int), ("tile", ctypes.c_int),
                ("psi", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("score", ctypes.c_float), ("mic", ctypes.c_int)]


class GsPeak(ctypes.Structure):
    _fields_ = [("mic", ctypes.c_int), ("tile", ctypes.c_int), ("view", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("cx", "cy", "psi", "dx", "dy", "defocus")]


class GsGrad(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("score", "d_psi", "d_dx", "d_dy", "d_defocus")]


class GsInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("l", "n_views", "n_psi", "nx", "ny", "tiles_x", "tiles_y", "box_x", "box_y",
                 "n_mics", "n_tiles")]


class GsTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("pick_total_ms", "micrograph_prep_ms", "corr_ms",
                                               "corr_pass2_ms", "corr_psi_ms")] + \
               [(n, ctypes.c_longlong) for n in ("corr_launches", "corr_sampled_maps", "corr_maps",
                                                  "refine_evals")]


class GsTune(ctypes.Structure):
    _fields_ = [("pass1_variant", ctypes.c_int), ("pass2_variant", ctypes.c_int),
                ("batch", ctypes.c_int), ("pipeline", ctypes.c_int),
                ("candidates", ctypes.c_int), ("tile_block", ctypes.c_int),
                ("psi_ms", ctypes.c_double),
                ("pass1_us", ctypes.c_double * 4), ("pass2_us", ctypes.c_double * 4)]


MAPSTAT_DTYPE = np.dtype([("score", "<f4"), ("pos", "<i4"), ("peak", "<f4"), ("sd", "<f4"),
                          ("sum", "<f8"), ("sum2", "<f8")])
PARTICLE_DTYPE = np.dtype([("psi_index", "<i4"), ("view", "<i4"), ("tile", "<i4"),
                           ("psi", "<f4"), ("cx", "<f4"), ("cy", "<f4"), ("score", "<f4"),
                           ("mic", "<i4")])

PEAK_DTYPE = np.dtype([("mic", "<i4"), ("tile", "<i4"), ("view", "<i4"), ("cx", "<f4"),
                       ("cy", "<f4"), ("psi", "<f4"), ("dx", "<f4"), ("dy", "<f4"),
                       ("defocus", "<f4")])
# gs_pick_rec: one line of a pick list for the duplicate removal
PICK_REC_DTYPE = np.dtype([("group", "<i4"), ("pad", "<i4"), ("x", "<f8"), ("y", "<f8"),
                           ("alt", "<f8"), ("az", "<f8"), ("phi", "<f8"), ("score", "<f8")])
GRAD_DTYPE = np.dtype([("score", "<f4"), ("d_psi", "<f4"), ("d_dx", "<f4"), ("d_dy", "<f4"),
                       ("d_defocus", "<f4")])


def psi_value(phi_step: float, index: int) -> float:
    """euler3 of in-plane step `index`: the reference's float accumulator
    (main.cu:1230)."""
    e = np.float32(0.0)
    for _ in range(index):
        e = np.float32(e + np.float32(phi_step))
    return float(e)


def peaks_from_picks(picks: np.ndarray, phi_step: float, defocus: float) -> np.ndarray:
    """Refinement candidates (PEAK_DTYPE) at the picked grid parameters."""
    out = np.zeros(len(picks), dtype=PEAK_DTYPE)
    for n in ("mic", "tile", "view", "cx", "cy"):
        out[n] = picks[n]
    out["psi"] = [psi_value(phi_step, int(i)) for i in picks["psi_index"]]
    out["defocus"] = defocus
    return out


GS_MODE_FUSED = 0
GS_MODE_PLAIN = 1
GS_PIPE_STREAM = 0
GS_PIPE_LAUNCH = 1
GS_PIPE_LANES = 2
GS_PIPE_SERIAL = 3

_lib = None


class GisspaError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = load().gs_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})")
        self.code = code


def load() -> ctypes.CDLL:
    """Load libgisspa.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C gisspa_amd` (or "
                           f"__graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.gs_params_default.argtypes = [ctypes.POINTER(GsParams)]
    lib.gs_params_default.restype = None
    lib.gs_strerror.argtypes = [ctypes.c_int]
    lib.gs_strerror.restype = ctypes.c_char_p
    lib.gs_create.argtypes = [ctypes.POINTER(GsParams), ctypes.c_int, ctypes.POINTER(vp)]
    lib.gs_set_templates.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(GsCtf)]
    lib.gs_set_micrograph.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(GsCtf)]
    lib.gs_set_micrograph_device.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(GsCtf)]
    lib.gs_set_micrographs.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(GsCtf), ctypes.c_int]
    lib.gs_set_mode.argtypes = [vp, ctypes.c_int]
    lib.gs_set_graphs.argtypes = [vp, ctypes.c_int]
    lib.gs_pick.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t,
                            ctypes.POINTER(ctypes.c_size_t)]
    lib.gs_get_picks.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.gs_get_mapstats.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.gs_get_template_real.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gs_get_tile_spectra.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gs_get_rotated_spectra.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    lib.gs_get_info.argtypes = [vp, ctypes.POINTER(GsInfo)]
    lib.gs_get_timing.argtypes = [vp, ctypes.POINTER(GsTiming)]
    lib.gs_set_batch.argtypes = [vp, ctypes.c_int]
    lib.gs_set_profiling.argtypes = [vp, ctypes.c_int]
    lib.gs_set_tile_block.argtypes = [vp, ctypes.c_int]
    lib.gs_set_record_capacity.argtypes = [vp, ctypes.c_longlong]
    lib.gs_set_pipeline.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.gs_adjoint.argtypes = [vp, vp, ctypes.c_int, vp]
    lib.gs_refine.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp]
    lib.gs_time_kernels.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int)]
    lib.gs_destroy.argtypes = [vp]
    lib.gs_remove_repeats.argtypes = [ctypes.c_int, vp, ctypes.c_size_t, ctypes.c_double,
                                      ctypes.c_double, vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.gs_remove_repeats_noalloc.argtypes = lib.gs_remove_repeats.argtypes
    lib.gs_project.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, vp]
    lib.gs_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(GsTune)]
    lib.gs_set_templates_mixed.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), vp,
                                           vp, ctypes.POINTER(GsCtf)]
    lib.gs_set_fft_variants.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.gs_io_error.argtypes = []
    lib.gs_io_error.restype = ctypes.c_char_p
    ip = ctypes.POINTER(ctypes.c_int)
    lib.gs_image_count.argtypes = [ctypes.c_char_p, ip]
    lib.gs_read_image.argtypes = [ctypes.c_char_p, ctypes.c_int, vp, ctypes.c_size_t, ip, ip, ip]
    lib.gs_hdf5_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, ctypes.c_size_t, ip,
                                 ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    lib.gs_hdf5_attr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, vp,
                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), vp,
                                 ctypes.c_size_t, ip]
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Entry points declared in include/gisspa.h."""
    hdr = os.path.join(_HERE, "..", "include", "gisspa.h")
    import re
    names = []
    with open(hdr) as f:
        for line in f:
            m = re.match(r"^\s*(?:int|void|const char\*)\s+(gs_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names


def pick_records(group, x, y, alt, az, phi, score) -> np.ndarray:
    """PICK_REC_DTYPE array from per-line columns."""
    n = len(x)
    r = np.zeros(n, dtype=PICK_REC_DTYPE)
    r["group"], r["x"], r["y"] = group, x, y
    r["alt"], r["az"], r["phi"], r["score"] = alt, az, phi, score
    return r


def remove_repeats(recs: np.ndarray, center: float, euler_thres: float,
                   device: int = 0) -> np.ndarray:
    """Keep mask of the duplicate removal (gs_remove_repeats; the result of
    python_scripts/remove_repeat_particles_from_list.py on the same lines)."""
    recs = np.ascontiguousarray(recs, dtype=PICK_REC_DTYPE)
    keep = np.zeros(len(recs), dtype=np.uint8)
    nk = ctypes.c_size_t()
    _check(load().gs_remove_repeats(device, recs.ctypes.data, len(recs), float(center),
                                    float(euler_thres), keep.ctypes.data, ctypes.byref(nk)),
           "gs_remove_repeats")
    return keep.astype(bool)


def project(vol: np.ndarray, eulers, device: int = 0) -> np.ndarray:
    """Projections [nviews][n][n] float32 of an n^3 volume [z][y][x] at
    (alt, az, phi) degrees (gs_project, EMAN convention)."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    n = v.shape[0]
    if v.shape != (n, n, n):
        raise ValueError("a cubic volume is needed")
    e = np.ascontiguousarray(np.asarray(eulers, dtype=np.float32).reshape(-1, 3))
    out = np.zeros((len(e), n, n), dtype=np.float32)
    _check(load().gs_project(device, v.ctypes.data, n, e.ctypes.data, len(e), out.ctypes.data),
           "gs_project")
    return out


def _io_check(code: int, what: str) -> None:
    if code == GS_ERR_IO:
        raise OSError(f"{what}: {load().gs_io_error().decode(errors='replace')}")
    _check(code, what)


def image_count(path: str) -> int:
    """Images in an MRC or EMAN HDF5 file (gs_image_count)."""
    n = ctypes.c_int()
    _io_check(load().gs_image_count(os.fsencode(path), ctypes.byref(n)), "gs_image_count")
    return n.value


def read_image(path: str, index: int = -1) -> np.ndarray:
    """Image `index` of an MRC or EMAN HDF5 file as float32 [ny][nx] (or
    [nz][ny][nx]); index < 0 reads the whole file (gs_read_image, the
    reference's emdata::readImage)."""
    lib = load()
    nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(path)
    _io_check(lib.gs_read_image(p, index, None, 0, ctypes.byref(nx), ctypes.byref(ny),
                                ctypes.byref(nz)), "gs_read_image")
    out = np.zeros((nz.value, ny.value, nx.value), np.float32)
    _io_check(lib.gs_read_image(p, index, out.ctypes.data, out.size, ctypes.byref(nx),
                                ctypes.byref(ny), ctypes.byref(nz)), "gs_read_image")
    return out[0] if nz.value == 1 else out


def hdf5_read(path: str, obj: str) -> np.ndarray:
    """An HDF5 dataset as float32 in its stored shape (gs_hdf5_read)."""
    lib = load()
    rank = ctypes.c_int()
    dims = (ctypes.c_longlong * 32)()
    p, o = os.fsencode(path), obj.encode()
    _io_check(lib.gs_hdf5_read(p, o, None, 0, ctypes.byref(rank), dims, 32), "gs_hdf5_read")
    out = np.zeros([dims[k] for k in range(rank.value)], np.float32)
    _io_check(lib.gs_hdf5_read(p, o, out.ctypes.data, out.size, ctypes.byref(rank), dims, 32),
              "gs_hdf5_read")
    return out


def hdf5_attr(path: str, obj: str, name: str):
    """An attribute of an HDF5 object: str, or a numpy float64 array of its
    numeric values (gs_hdf5_attr)."""
    lib = load()
    num = np.zeros(4096, np.float64)
    s = ctypes.create_string_buffer(1 << 16)
    n = ctypes.c_size_t()
    is_str = ctypes.c_int()
    _io_check(lib.gs_hdf5_attr(os.fsencode(path), obj.encode(), name.encode(), num.ctypes.data,
                               num.size, ctypes.byref(n), s, len(s), ctypes.byref(is_str)),
              "gs_hdf5_attr")
    return s.value.decode() if is_str.value else num[:n.value].copy()


def make_params(**kw) -> GsParams:
    """gs_params from config-key style names (angpix, phistep, Highres, ...)."""
    p = GsParams()
    load().gs_params_default(ctypes.byref(p))
    alias = {"angpix": "apix", "phistep": "phi_step", "Highres": "highres", "Lowres": "lowres",
             "threshold": "threshold"}
    for k, v in kw.items():
        name = alias.get(k, k)
        if name in ("window_size", "phase_flip", "overlap", "norm_type", "invert"):
            setattr(p, name, int(v))
        elif hasattr(p, name):
            setattr(p, name, float(v))
    return p


def _check(code: int, what: str) -> None:
    if code != GS_OK:
        raise GisspaError(code, what)


class Picker:
    """One picking context on one device (gs_create ... gs_destroy)."""

    def __init__(self, device: int = 0, **params):
        self.lib = load()
        self.params = make_params(**params)
        self.ctx = ctypes.c_void_p()
        _check(self.lib.gs_create(ctypes.byref(self.params), device, ctypes.byref(self.ctx)),
               "gs_create")

    def close(self) -> None:
        if self.ctx:
            self.lib.gs_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_templates(self, tm: np.ndarray, defocus: float, dfdiff: float = 0.0,
                      dfang: float = 0.0) -> None:
        tm = np.ascontiguousarray(tm, dtype=np.float32)
        n, by, bx = tm.shape
        c = GsCtf(defocus, dfdiff, dfang)
        _check(self.lib.gs_set_templates(self.ctx, tm.ctypes.data, n, bx, by, ctypes.byref(c)),
               "gs_set_templates")

    def set_templates_mixed(self, templates, defocus: float, dfdiff: float = 0.0,
                            dfang: float = 0.0) -> None:
        """Templates of different boxes (a list of 2D arrays)."""
        arrs = [np.ascontiguousarray(t, dtype=np.float32) for t in templates]
        n = len(arrs)
        ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in arrs])
        bx = np.array([a.shape[1] for a in arrs], dtype=np.int32)
        by = np.array([a.shape[0] for a in arrs], dtype=np.int32)
        c = GsCtf(defocus, dfdiff, dfang)
        _check(self.lib.gs_set_templates_mixed(self.ctx, n, ptrs, bx.ctypes.data, by.ctypes.data,
                                               ctypes.byref(c)), "gs_set_templates_mixed")

    def set_micrograph(self, img: np.ndarray, defocus: float, dfdiff: float = 0.0,
                       dfang: float = 0.0) -> None:
        img = np.ascontiguousarray(img, dtype=np.float32)
        ny, nx = img.shape
        c = GsCtf(defocus, dfdiff, dfang)
        _check(self.lib.gs_set_micrograph(self.ctx, img.ctypes.data, nx, ny, ctypes.byref(c)),
               "gs_set_micrograph")

    def set_micrograph_device(self, ptr: int, nx: int, ny: int, defocus: float,
                              dfdiff: float = 0.0, dfang: float = 0.0) -> None:
        c = GsCtf(defocus, dfdiff, dfang)
        _check(self.lib.gs_set_micrograph_device(self.ctx, ctypes.c_void_p(ptr), nx, ny,
                                                 ctypes.byref(c)), "gs_set_micrograph_device")

    def set_micrographs(self, imgs, ctfs) -> None:
        """A batch of micrographs (any sizes); ctfs: one (defocus, dfdiff, dfang)
        tuple or defocus per micrograph.  Picks then carry `mic`."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in imgs]
        n = len(arrs)
        ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in arrs])
        nx = (ctypes.c_int * n)(*[a.shape[1] for a in arrs])
        ny = (ctypes.c_int * n)(*[a.shape[0] for a in arrs])
        cs = (GsCtf * n)()
        for i, c in enumerate(ctfs):
            c = tuple(np.atleast_1d(c)) + (0.0, 0.0)
            cs[i] = GsCtf(float(c[0]), float(c[1]), float(c[2]))
        _check(self.lib.gs_set_micrographs(self.ctx, n, ptrs, nx, ny, cs, 0), "gs_set_micrographs")

    def set_mode(self, mode: int) -> None:
        _check(self.lib.gs_set_mode(self.ctx, mode), "gs_set_mode")

    def set_graphs(self, enable: bool) -> None:
        _check(self.lib.gs_set_graphs(self.ctx, int(bool(enable))), "gs_set_graphs")

    def pick(self, psi_begin: int = 0, psi_end: int = -1) -> np.ndarray:
        """Picks of in-plane steps [psi_begin, psi_end) (PARTICLE_DTYPE records)."""
        n = ctypes.c_size_t(0)
        r = self.lib.gs_pick(self.ctx, psi_begin, psi_end, None, 0, ctypes.byref(n))
        if r != GS_ERR_CAPACITY:
            _check(r, "gs_pick")
        out = np.zeros(n.value, dtype=PARTICLE_DTYPE)
        if n.value:
            _check(self.lib.gs_get_picks(self.ctx, out.ctypes.data, n.value, ctypes.byref(n)),
                   "gs_get_picks")
        return out

    def mapstats(self) -> np.ndarray:
        n = ctypes.c_size_t(0)
        self.lib.gs_get_mapstats(self.ctx, None, 0, ctypes.byref(n))
        out = np.zeros(n.value, dtype=MAPSTAT_DTYPE)
        _check(self.lib.gs_get_mapstats(self.ctx, out.ctypes.data, n.value, ctypes.byref(n)),
               "gs_get_mapstats")
        return out

    def info(self) -> GsInfo:
        i = GsInfo()
        _check(self.lib.gs_get_info(self.ctx, ctypes.byref(i)), "gs_get_info")
        return i

    def timing(self) -> GsTiming:
        t = GsTiming()
        _check(self.lib.gs_get_timing(self.ctx, ctypes.byref(t)), "gs_get_timing")
        return t

    def set_profiling(self, enable: bool) -> None:
        _check(self.lib.gs_set_profiling(self.ctx, int(bool(enable))), "gs_set_profiling")

    def adjoint(self, peaks: np.ndarray) -> np.ndarray:
        """Refinement objective and gradient (GRAD_DTYPE) of PEAK_DTYPE candidates."""
        pk = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE)
        out = np.zeros(len(pk), dtype=GRAD_DTYPE)
        _check(self.lib.gs_adjoint(self.ctx, pk.ctypes.data, len(pk), out.ctypes.data),
               "gs_adjoint")
        return out

    def refine(self, peaks: np.ndarray, iters: int = 20):
        """Refined candidates and their final (objective, gradient)."""
        pk = np.array(peaks, dtype=PEAK_DTYPE, copy=True)
        out = np.zeros(len(pk), dtype=GRAD_DTYPE)
        _check(self.lib.gs_refine(self.ctx, pk.ctypes.data, len(pk), iters, out.ctypes.data),
               "gs_refine")
        return pk, out

    def set_pipeline(self, mode: int, lag: int = 0, ring: int = 0) -> None:
        _check(self.lib.gs_set_pipeline(self.ctx, mode, lag, ring), "gs_set_pipeline")

    def set_tile_block(self, n: int) -> None:
        _check(self.lib.gs_set_tile_block(self.ctx, n), "gs_set_tile_block")

    def set_record_capacity(self, records: int) -> None:
        """Initial capacity of the device-side pick buffers (gs_set_record_capacity)."""
        _check(self.lib.gs_set_record_capacity(self.ctx, records), "gs_set_record_capacity")

    def autotune(self, psi_index: int = 0) -> GsTune:
        """Time the correlation schedules on the device and keep the fastest
        (gs_autotune)."""
        t = GsTune()
        _check(self.lib.gs_autotune(self.ctx, psi_index, ctypes.byref(t)), "gs_autotune")
        return t

    def set_fft_variants(self, pass1: int, pass2: int) -> None:
        _check(self.lib.gs_set_fft_variants(self.ctx, pass1, pass2), "gs_set_fft_variants")

    def set_batch(self, n: int) -> None:
        _check(self.lib.gs_set_batch(self.ctx, n), "gs_set_batch")

    def time_kernels(self, psi_index: int = 0, reps: int = 10):
        """(pass1_ms, pass2_ms, maps_per_launch): mean device time of one launch."""
        ms = np.zeros(2, dtype=np.float64)
        n = ctypes.c_int(0)
        _check(self.lib.gs_time_kernels(self.ctx, psi_index, reps, ms.ctypes.data, ctypes.byref(n)),
               "gs_time_kernels")
        return float(ms[0]), float(ms[1]), n.value

    def template_real(self) -> np.ndarray:
        i = self.info()
        out = np.zeros((i.n_views, i.l, i.l), dtype=np.float32)
        _check(self.lib.gs_get_template_real(self.ctx, out.ctypes.data, out.size),
               "gs_get_template_real")
        return out

    def tile_spectra(self) -> np.ndarray:
        """[tile][kx][ky] complex64 half spectra."""
        i = self.info()
        out = np.zeros((i.n_tiles, i.l // 2 + 1, i.l), dtype=np.complex64)
        _check(self.lib.gs_get_tile_spectra(self.ctx, out.ctypes.data, out.size * 2),
               "gs_get_tile_spectra")
        return out

    def rotated_spectra(self, psi_index: int) -> np.ndarray:
        i = self.info()
        out = np.zeros((i.n_views, i.l // 2 + 1, i.l), dtype=np.complex64)
        _check(self.lib.gs_get_rotated_spectra(self.ctx, psi_index, out.ctypes.data, out.size * 2),
               "gs_get_rotated_spectra")
        return out
