"""ctypes view of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The picker never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("apix", "phi_step", "kk", "energy", "cs", "highres", "lowres", "diameter",
                 "threshold", "ampconst", "a", "b", "b2", "bfactor", "bfactor2", "bfactor3")] + \
               [(n, ctypes.c_int) for n in
                ("l", "l_config", "overlap", "phase_flip", "invert", "norm_type")]


class OrCtf(ctypes.Structure):
    _fields_ = [("defocus_um", ctypes.c_double), ("dfdiff", ctypes.c_double),
                ("dfang", ctypes.c_double)]


MAPSTAT = np.dtype([("peak", "<f8"), ("sum", "<f8"), ("sum2", "<f8"), ("score", "<f8"),
                    ("pos", "<i4"), ("pad", "<i4")])
PICK1 = np.dtype([("psi_index", "<i4"), ("view", "<i4"), ("cx", "<f8"), ("cy", "<f8"),
                  ("score", "<f8")])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        P = ctypes.POINTER(OrParams)
        C = ctypes.POINTER(OrCtf)
        lib.or_fft1d.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        lib.or_fft2d.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.or_ctf_ast.argtypes = [ctypes.c_int] * 4 + [P, C, ctypes.c_int]
        lib.or_ctf_ast.restype = ctypes.c_double
        lib.or_tile_offsets.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        lib.or_template_prep_ctf.argtypes = [P, C, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        lib.or_rotate_normalize.argtypes = [P, vp, ctypes.c_int, ctypes.c_float, vp]
        lib.or_micrograph_prep.argtypes = [P, C, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        lib.or_correlate_norm0.argtypes = [P, vp, vp, vp]
        lib.or_cc_map.argtypes = [P, vp, vp, vp]
        lib.or_correlate_many.argtypes = [P, vp, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int]
        lib.or_rotate_bilinear.argtypes = [vp, ctypes.c_int, ctypes.c_float, vp]
        lib.or_pick_norm0.argtypes = [P, C, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                      ctypes.c_int]
        lib.or_pick_norm1.argtypes = [P, C, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                      ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int]
        lib.or_select_norm0.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        lib.or_num_psi.argtypes = [ctypes.c_double]
        lib.or_template_whitened.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        lib.or_refine_score.argtypes = [P, C, vp, vp, ctypes.c_double, ctypes.c_double, vp, vp]
        lib.or_psi_value.argtypes = [ctypes.c_double, ctypes.c_int]
        lib.or_psi_value.restype = ctypes.c_float
        d = ctypes.c_double
        lib.or_euler_distance.argtypes = [d] * 6 + [vp]
        lib.or_remove_repeats.argtypes = [ctypes.c_int] + [vp] * 7 + [d, d, vp]
        lib.or_project.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp]
        lib.or_set_threads.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def params(l: int, **kw) -> OrParams:
    """Oracle parameters from config-key names, with the reference defaults."""
    p = OrParams()
    d = dict(apix=-1, phi_step=-1, kk=-1, energy=-1, cs=-1, highres=-1, lowres=-1, diameter=-1,
             threshold=7, ampconst=0.07, a=-9.32, b=2.65, b2=0.01908, bfactor=-78.7757,
             bfactor2=-12.9121, bfactor3=1.28732, overlap=-1, phase_flip=0, invert=0, norm_type=0)
    alias = {"angpix": "apix", "phistep": "phi_step", "Highres": "highres", "Lowres": "lowres"}
    for k, v in kw.items():
        k = alias.get(k, k)
        if k in d:
            d[k] = v
    # the reference's fixed constants are float (DataReader.h:23-25)
    for k in ("ampconst", "a", "b", "b2", "bfactor", "bfactor2", "bfactor3"):
        d[k] = float(np.float32(d[k]))
    for k, v in d.items():
        if isinstance(getattr(p, k), float):
            v = float(np.float32(v))  # the reference stores every parameter as float
        setattr(p, k, type(getattr(p, k))(v))
    p.l = l
    p.l_config = kw.get("window_size", l)
    return p


def ctf(defocus: float, dfdiff: float = 0.0, dfang: float = 0.0) -> OrCtf:
    return OrCtf(defocus, dfdiff, dfang)


def fft2(a: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.complex128).copy()
    ny, nx = x.shape
    load().or_fft2d(x.ctypes.data, nx, ny, sign)
    return x


def template_prep(p: OrParams, c: OrCtf, tm: np.ndarray) -> np.ndarray:
    tm = np.ascontiguousarray(tm, dtype=np.float32)
    n, by, bx = tm.shape
    out = np.zeros((n, p.l, p.l), dtype=np.complex128)
    r = load().or_template_prep_ctf(ctypes.byref(p), ctypes.byref(c), tm.ctypes.data, n, bx, by,
                                    out.ctypes.data)
    assert r == 0, r
    return out


def rotate_normalize(p: OrParams, treal: np.ndarray, euler3: float) -> np.ndarray:
    treal = np.ascontiguousarray(treal, dtype=np.complex128)
    out = np.zeros_like(treal)
    load().or_rotate_normalize(ctypes.byref(p), treal.ctypes.data, treal.shape[0], euler3,
                               out.ctypes.data)
    return out


def tile_offsets(n: int, l: int, overlap: int) -> list[int]:
    buf = np.zeros(8192, dtype=np.int32)
    k = load().or_tile_offsets(n, l, overlap, buf.ctypes.data, 8192)
    return list(buf[:k])


def micrograph_prep(p: OrParams, c: OrCtf, img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float32)
    ny, nx = img.shape
    T = len(tile_offsets(nx, p.l, p.overlap)) * len(tile_offsets(ny, p.l, p.overlap))
    out = np.zeros((T, p.l, p.l), dtype=np.complex128)
    r = load().or_micrograph_prep(ctypes.byref(p), ctypes.byref(c), img.ctypes.data, nx, ny,
                                  out.ctypes.data, T)
    assert r == T, r
    return out


def template_whitened(p: OrParams, tm: np.ndarray) -> np.ndarray:
    """Whitened template spectra [n][l][l] (before the CTF weighting)."""
    tm = np.ascontiguousarray(tm, dtype=np.float32)
    n, by, bx = tm.shape
    out = np.zeros((n, p.l, p.l), dtype=np.complex128)
    r = load().or_template_whitened(ctypes.byref(p), tm.ctypes.data, n, bx, by, out.ctypes.data)
    assert r == 0, r
    return out


def refine_score(p: OrParams, c0: OrCtf, white: np.ndarray, tile: np.ndarray, x0: float,
                 y0: float, theta) -> tuple[float, float, float]:
    """Refinement objective of one candidate: (score, c, sigma) at
    theta = (psi degrees, dx, dy, defocus um)."""
    w = np.ascontiguousarray(white, dtype=np.complex128)
    t = np.ascontiguousarray(tile, dtype=np.complex128)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    load().or_refine_score(ctypes.byref(p), ctypes.byref(c0), w.ctypes.data, t.ctypes.data,
                           float(x0), float(y0), th.ctypes.data, out.ctypes.data)
    return float(out[0]), float(out[1]), float(out[2])


def refine_grad_fd(p: OrParams, c0: OrCtf, white, tile, x0, y0, theta, h=1e-3) -> np.ndarray:
    """FP64 central differences of refine_score in each of the 4 parameters."""
    g = np.zeros(4)
    for i in range(4):
        tp = np.array(theta, dtype=np.float64)
        tm_ = tp.copy()
        tp[i] += h
        tm_[i] -= h
        g[i] = (refine_score(p, c0, white, tile, x0, y0, tp)[0] -
                refine_score(p, c0, white, tile, x0, y0, tm_)[0]) / (2 * h)
    return g


def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (whole-image FFTs, tiles)."""
    load().or_set_threads(int(n))


def num_psi(phi_step: float) -> int:
    return load().or_num_psi(phi_step)


def psi_value(phi_step: float, i: int) -> float:
    return load().or_psi_value(phi_step, i)


def pick_norm0(p: OrParams, c: OrCtf, tm: np.ndarray, img: np.ndarray, psi_begin: int = 0,
               psi_end: int | None = None, nthreads: int = 0) -> np.ndarray:
    """Per-map results [psi][tile][view] (MAPSTAT records)."""
    tm = np.ascontiguousarray(tm, dtype=np.float32)
    img = np.ascontiguousarray(img, dtype=np.float32)
    n, by, bx = tm.shape
    ny, nx = img.shape
    if psi_end is None:
        psi_end = num_psi(p.phi_step)
    T = len(tile_offsets(nx, p.l, p.overlap)) * len(tile_offsets(ny, p.l, p.overlap))
    out = np.zeros((psi_end - psi_begin, T, n), dtype=MAPSTAT)
    r = load().or_pick_norm0(ctypes.byref(p), ctypes.byref(c), tm.ctypes.data, n, bx, by,
                             img.ctypes.data, nx, ny, psi_begin, psi_end, out.ctypes.data, nthreads)
    assert r == T, r
    return out


def select_norm0(p: OrParams, st: np.ndarray, psi_begin: int, nx: int, ny: int) -> np.ndarray:
    st = np.ascontiguousarray(st)
    npsi, T, n = st.shape
    cap = st.size
    out = np.zeros(max(cap, 1), dtype=PICK1)
    k = load().or_select_norm0(ctypes.byref(p), st.ctypes.data, npsi, psi_begin, n, nx, ny,
                               out.ctypes.data, cap)
    assert k >= 0
    return out[:k]


def pick_norm1(p: OrParams, c: OrCtf, tm: np.ndarray, img: np.ndarray, cap: int = 1 << 20,
               nthreads: int = 0) -> np.ndarray:
    tm = np.ascontiguousarray(tm, dtype=np.float32)
    img = np.ascontiguousarray(img, dtype=np.float32)
    n, by, bx = tm.shape
    ny, nx = img.shape
    out = np.zeros(cap, dtype=PICK1)
    k = load().or_pick_norm1(ctypes.byref(p), ctypes.byref(c), tm.ctypes.data, n, bx, by,
                             img.ctypes.data, nx, ny, out.ctypes.data, cap, nthreads)
    assert k >= 0
    return out[:min(k, cap)]


def cc_map(p: OrParams, tile_spec: np.ndarray, tspec: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(tile_spec, dtype=np.complex128)
    b = np.ascontiguousarray(tspec, dtype=np.complex128)
    out = np.zeros((p.l, p.l), dtype=np.float64)
    load().or_cc_map(ctypes.byref(p), a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def correlate_norm0(p: OrParams, tile_spec: np.ndarray, tspec: np.ndarray):
    a = np.ascontiguousarray(tile_spec, dtype=np.complex128)
    b = np.ascontiguousarray(tspec, dtype=np.complex128)
    out = np.zeros(1, dtype=MAPSTAT)
    load().or_correlate_norm0(ctypes.byref(p), a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out[0]


def rotate_bilinear(img: np.ndarray, e: float) -> np.ndarray:
    a = np.ascontiguousarray(img, dtype=np.float64)
    out = np.zeros_like(a)
    load().or_rotate_bilinear(a.ctypes.data, a.shape[0], e, out.ctypes.data)
    return out


def ctf_ast(p: OrParams, c: OrCtf, x: int, y: int, nx: int, ny: int, mode: int) -> float:
    return load().or_ctf_ast(x, y, nx, ny, ctypes.byref(p), ctypes.byref(c), mode)


def correlate_many(p: OrParams, tile_spec: np.ndarray, tspec: np.ndarray, nthreads: int = 0):
    a = np.ascontiguousarray(tile_spec, dtype=np.complex128)
    b = np.ascontiguousarray(tspec, dtype=np.complex128)
    out = np.zeros((a.shape[0], b.shape[0]), dtype=MAPSTAT)
    load().or_correlate_many(ctypes.byref(p), a.ctypes.data, a.shape[0], b.ctypes.data,
                             b.shape[0], out.ctypes.data, nthreads)
    return out


def euler_distance(e1, e2):
    """The duplicate-removal script's orientation distance in degrees between
    outer pick e1 = (alt, az, phi) and inner pick e2, or None where the script
    skips the pair (remove_repeat_particles_from_list.py:83-132)."""
    out = ctypes.c_double()
    ok = load().or_euler_distance(*(float(v) for v in e1), *(float(v) for v in e2),
                                  ctypes.byref(out))
    return out.value if ok else None


def remove_repeats(group, x, y, alt, az, phi, score, center: float, euler_thres: float):
    """Keep mask of remove_repeat_particles_from_list.py (True = line kept)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, alt, az, phi, score)]
    g = np.ascontiguousarray(group, dtype=np.int32)
    n = len(g)
    keep = np.zeros(n, dtype=np.uint8)
    r = load().or_remove_repeats(n, g.ctypes.data, *(a.ctypes.data for a in arrs),
                                 float(center), float(euler_thres), keep.ctypes.data)
    assert r == 0, r
    return keep.astype(bool)


def project(vol: np.ndarray, eulers) -> np.ndarray:
    """Projections [nviews][n][n] (FP64) of an n^3 volume [z][y][x] at
    (alt, az, phi) degrees, EMAN convention."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    n = v.shape[0]
    assert v.shape == (n, n, n)
    e = np.ascontiguousarray(np.asarray(eulers, dtype=np.float64).reshape(-1, 3))
    out = np.zeros((len(e), n, n), dtype=np.float64)
    r = load().or_project(v.ctypes.data, n, e.ctypes.data, len(e), out.ctypes.data)
    assert r == 0, r
    return out
