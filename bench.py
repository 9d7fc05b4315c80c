#!/usr/bin/env python3
"""Benchmark: micrographs/hour of the GisSPA picker on one MI355X.

Workload (the full-size plain-method configuration): one 5760x4092 float32
micrograph with 300 embedded particles (SNR 0.05), 180x180 templates of a
60-blob synthetic density at 1649 views of a ~5 degree sphere, 72 in-plane
steps (5 degrees), defocus 1.8 um, window 720 with overlap 180 (88 tiles).
One step = one micrograph through the whole hot path: micrograph prep (stats,
hot-pixel clamp, mean removal, 88 tile FFTs) and, for every psi, template
rotation + FFT and all 88 x 1649 cross-correlation maps with their peak /
score reduction and pick selection.  The micrograph is resident in device
memory when the timed region starts.

Prints one JSON line (rank 0).  --gpus N with N > 1 exits: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per CC map at window 720 from rocprofv3 FETCH_SIZE (x2.0) and
# WRITE_SIZE (x1.0) counters, scaled by a 1 GiB copy kernel's readings
# (profiles/corr_720_full_workload.md)
TRAFFIC_PER_MAP_720 = {"k_corr_pass1": 4.084e6, "k_corr_pass2": 1.933e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime access for device-resident inputs."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")

    def malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        r = self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes))
        if r != 0:
            raise RuntimeError(f"hipMalloc failed: {r}")
        return p.value

    def h2d(self, dst: int, arr: np.ndarray) -> None:
        r = self.lib.hipMemcpy(ctypes.c_void_p(dst), arr.ctypes.data_as(ctypes.c_void_p),
                               ctypes.c_size_t(arr.nbytes), 1)
        if r != 0:
            raise RuntimeError(f"hipMemcpy failed: {r}")

    def d2d(self, dst: int, src: int, nbytes: int) -> None:
        r = self.lib.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), 3)
        if r != 0:
            raise RuntimeError(f"hipMemcpy failed: {r}")

    def sync(self) -> None:
        self.lib.hipDeviceSynchronize()

    def copy_rate(self, nbytes: int, reps: int = 3) -> float:
        """Best device-to-device copy rate (bytes read + written per second)
        of an nbytes buffer: what the HBM sustains for a plain stream."""
        a, b = self.malloc(nbytes), self.malloc(nbytes)
        best = 0.0
        try:
            self.d2d(b, a, nbytes)
            for _ in range(reps):
                self.sync()
                t = time.perf_counter()
                self.d2d(b, a, nbytes)
                self.sync()
                best = max(best, 2.0 * nbytes / (time.perf_counter() - t))
        finally:
            self.free(a)
            self.free(b)
        return best

    def free(self, p: int) -> None:
        self.lib.hipFree(ctypes.c_void_p(p))


REF_EULERS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                          "euler_C1_step5.npy")


def reference_views() -> np.ndarray:
    """(alt, az) of the 1649 views of the reference's Euler_file/euler_C1_step5.lst."""
    return np.load(REF_EULERS)[:, :2].astype(np.float64)


def full_workload(cache_dir: str):
    from gisspa_amd import synth
    path = os.path.join(cache_dir, "gisspa_full_c1step5_seed0.npz")
    if os.path.exists(path):
        d = np.load(path)
        return d["templates"], d["image"], float(d["defocus"])
    t = time.time()
    wl = synth.full_workload(seed=0, eulers=reference_views())
    img, df = wl.micrographs[0]
    log(f"[bench] generated full workload in {time.time() - t:.1f}s")
    try:
        os.makedirs(cache_dir, exist_ok=True)
        np.savez(path, templates=wl.templates, image=img, defocus=df)
    except OSError:
        pass
    return wl.templates, img, df


FULL_PARAMS = dict(angpix=1.0, phistep=5.0, kk=3, energy=300, cs=2.7, Highres=8, Lowres=100,
                   diameter=135.0, threshold=7.0, window_size=720, overlap=180, ampconst=0.1)


def cpu_baseline(templates, image, defocus, nthreads: int, n_views: int):
    """FP64 oracle timed on a bounded sample of the same workload: the whole
    micrograph preparation (88 tile spectra), then for one psi the rotation of
    n_views templates and their correlation with all 88 tiles.  Per-micrograph
    time = prep + 72 psi x (N / n_views) x (rotation + correlation of the
    sample); template preparation is once per run and excluded on both sides."""
    from oracle import oracle as orc
    p = orc.params(720, **FULL_PARAMS)
    c = orc.ctf(defocus)
    orc.set_threads(nthreads)
    t0 = time.time()
    tiles = orc.micrograph_prep(p, c, image)
    t_prep = time.time() - t0
    orc.set_threads(1)
    t0 = time.time()
    orc.micrograph_prep(p, c, image)
    t_prep1 = time.time() - t0
    treal = orc.template_prep(p, c, np.ascontiguousarray(templates[:n_views]))
    t0 = time.time()
    spec = orc.rotate_normalize(p, treal, 0.0)
    t_rot = time.time() - t0
    t0 = time.time()
    orc.correlate_many(p, tiles, spec, nthreads)
    t_maps = time.time() - t0
    # the same maps on one thread: one view against all tiles
    t0 = time.time()
    orc.correlate_many(p, tiles, spec[:1], 1)
    t_map1 = (time.time() - t0) / tiles.shape[0]
    orc.set_threads(nthreads)
    T = tiles.shape[0]
    N = templates.shape[0]
    sec = t_prep + 72 * (N / n_views) * (t_rot / nthreads + t_maps)
    sec1 = t_prep1 + 72 * N * (t_rot / n_views + T * t_map1)
    return dict(value=3600.0 / sec, unit="micrographs/hour", cores=nthreads, kind="port",
                single_thread_value=3600.0 / sec1,
                sample=f"FP64 oracle on the host: micrograph prep {t_prep:.1f}s ({nthreads} "
                       f"threads; {t_prep1:.1f}s on 1), rotation of {n_views} views "
                       f"{t_rot:.2f}s (1 thread, divided by {nthreads} for the threaded rate), "
                       f"{n_views} views x {T} tiles = {n_views * T} CC maps {t_maps:.1f}s "
                       f"({nthreads} threads), 1 view x {T} tiles {T * t_map1:.1f}s "
                       f"(1 thread); extrapolated to {N} views x 72 psi")


def plain_rate(templates, d_img, nx, ny, defocus, npsi: int):
    """Stage-by-stage picker (GS_MODE_PLAIN) on the same micrograph: prep +
    npsi in-plane steps, timed on the wall clock, extrapolated to 72 steps."""
    import gisspa_amd
    pk = gisspa_amd.Picker(0, **FULL_PARAMS)
    pk.set_mode(gisspa_amd.GS_MODE_PLAIN)
    pk.set_templates(templates, defocus)
    pk.set_micrograph_device(d_img, nx, ny, defocus)   # warm-up of every kernel
    pk.pick(0, 1)
    t0 = time.time()
    pk.set_micrograph_device(d_img, nx, ny, defocus)
    t_prep = time.time() - t0
    t0 = time.time()
    pk.pick(0, npsi)
    t_psi = (time.time() - t0) / npsi
    pk.close()
    sec = t_prep + 72 * t_psi
    return dict(value=3600.0 / sec, unit="micrographs/hour",
                sample=f"GS_MODE_PLAIN: micrograph prep {t_prep:.2f}s + {npsi} psi at "
                       f"{t_psi:.2f}s each, extrapolated to 72 psi")


def small_config():
    """The small plain-method configuration run in full on the GPU: one
    1024^2 micrograph, 64 views of a 128^3 density, window 512 / overlap 128
    (9 tiles), 36 psi = 20,736 CC maps; wall clock of micrograph preparation
    and the whole pick, after one warm-up pick.  (cpu_baseline times the
    FP64 oracle on the same maps and compares the pick lists.)"""
    import gisspa_amd
    from gisspa_amd import synth
    wl = synth.small_workload(seed=0)
    img, df = wl.micrographs[0]
    pk = gisspa_amd.Picker(0, **wl.params)
    pk.set_templates(wl.templates, df)
    pk.set_micrograph(img, df)
    pk.pick()
    t0 = time.perf_counter()
    pk.set_micrograph(img, df)
    picks = pk.pick()
    t_gpu = time.perf_counter() - t0
    pk.close()
    out = {"maps": 9 * 64 * 36, "gpu_s": t_gpu, "gpu_micrographs_per_hour": 3600.0 / t_gpu,
           "picks": int(len(picks)), "cpu_s": None, "cpu_threads": None, "identical_picks": None}
    return out, wl, picks


def cpu_baseline_small(wl, picks, nthreads: int):
    """The FP64 oracle on the small configuration's 20,736 maps (micrograph
    and template preparation included) on nthreads threads, and whether its
    pick list equals the GPU's (psi, view, centre)."""
    from oracle import oracle as orc
    img, df = wl.micrographs[0]
    p = orc.params(512, **wl.params)
    t0 = time.perf_counter()
    st = orc.pick_norm0(p, orc.ctf(df), wl.templates, img, nthreads=nthreads)
    sel = orc.select_norm0(p, st, 0, img.shape[1], img.shape[0])
    t = time.perf_counter() - t0
    same = bool(len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"]) and
                np.array_equal(sel["psi_index"], picks["psi_index"]) and
                np.array_equal(sel["cx"].astype(np.float32), picks["cx"]) and
                np.array_equal(sel["cy"].astype(np.float32), picks["cy"]))
    return t, same


def prep_times(templates, d_img, nx, ny, defocus, hip):
    """Micrograph preparation alone (statistics, hot-pixel clamp, tiles and
    their spectra; with phase_flip 1 also the whole-micrograph CTF flip,
    whitening and band pass, main.cu:523-655), device synchronised, best of
    three, for phase_flip 0 and 1."""
    import gisspa_amd
    out = {}
    for pf in (0, 1):
        pk = gisspa_amd.Picker(0, **dict(FULL_PARAMS, phase_flip=pf))
        pk.set_templates(np.ascontiguousarray(templates[:8]), defocus)
        best = 1e30
        for _ in range(3):
            hip.sync()
            t0 = time.perf_counter()
            pk.set_micrograph_device(d_img, nx, ny, defocus)
            hip.sync()
            best = min(best, time.perf_counter() - t0)
        pk.close()
        out[f"phase_flip_{pf}_ms"] = best * 1e3
    return out


def norm1_rate(templates, d_img, nx, ny, defocus, n_views: int):
    """norm_type 1 (per-pixel normalisation over all views x psi,
    main.cu:1170-1229) on the same micrograph with the first n_views views:
    both passes over all 72 psi, wall clock, scaled to all views (the work is
    linear in the number of views).  The threshold is the reference's default
    (7): z-scores of the per-pixel normalisation are not on the scale of the
    norm_type 0 scores the headline run's threshold is set for."""
    import gisspa_amd
    pk = gisspa_amd.Picker(0, **dict(FULL_PARAMS, norm_type=1, threshold=7.0))
    pk.set_templates(np.ascontiguousarray(templates[:n_views]), defocus)
    pk.set_micrograph_device(d_img, nx, ny, defocus)
    t0 = time.time()
    picks = pk.pick()
    t = time.time() - t0
    pk.close()
    sec = t * templates.shape[0] / n_views
    return dict(value=3600.0 / sec, unit="micrographs/hour", views=n_views, seconds=t,
                picks=int(len(picks)),
                sample=f"norm_type 1, {n_views} of {templates.shape[0]} views x 72 psi (both "
                       f"passes) in {t:.2f}s, scaled by the view count")


def io_rate(templates, image, cache_dir: str, n_lines: int = 100000, seed: int = 5):
    """Host-side formats around the path: reading the template stack and the
    micrograph through gs_read_image as MRC and as EMAN2 HDF5 (contiguous;
    chunked + deflate), bit-checked against the arrays written; the score
    filter and the lst -> RELION STAR conversion of a synthetic pick list of
    n_lines lines."""
    import gisspa_amd
    from gisspa_amd import lstconv, mrc
    from tests import _hdf5 as H
    d = os.path.join(cache_dir, "gisspa_io_bench")
    os.makedirs(d, exist_ok=True)
    out = {"reads": {}}
    cases = [
        ("templates, MRC stack", "t.mrcs", lambda p: mrc.write_mrc(p, templates), templates),
        ("templates, EMAN2 HDF5 contiguous", "t.hdf",
         lambda p: H.Writer(0).write(p, H.eman2_stack(list(templates))), templates),
        ("templates, EMAN2 HDF5 deflate", "tz.hdf",
         lambda p: H.Writer(0).write(p, H.eman2_stack(list(templates), layout="chunked",
                                                      chunks=templates.shape[1:], deflate=1)),
         templates),
        ("micrograph, MRC", "m.mrc", lambda p: mrc.write_mrc(p, image[None]), image),
        ("micrograph, EMAN1 HDF5", "m.hdf", lambda p: H.Writer(0).write(p, H.eman1_stack([image])),
         image),
    ]
    for name, fn, write, ref in cases:
        p = os.path.join(d, fn)
        write(p)
        t0 = time.time()
        got = gisspa_amd.read_image(p, -1)
        t = time.time() - t0
        out["reads"][name] = {"seconds": t, "MB_per_s": ref.nbytes / 1e6 / t,
                              "identical": bool(np.array_equal(got, ref))}
        log(f"[bench] io: {name}: {ref.nbytes / 1e6:.0f} MB in {t * 1e3:.0f} ms "
            f"({ref.nbytes / 1e6 / t:.0f} MB/s), identical {out['reads'][name]['identical']}")
    rng = np.random.default_rng(seed)
    lines = ["%d\t/data/mic%03d.mrc\tdefocus=%f\tdfdiff=%f\tdfang=%f\teuler=%f,%f,%f\t"
             "center=%f,%f\tscore=%f\n" % (i % 300, i // 300, rng.uniform(1, 3), 0.0, 0.0,
                                             rng.uniform(0, 180), rng.uniform(0, 360),
                                             rng.uniform(0, 360), rng.uniform(0, 5760),
                                             rng.uniform(0, 4092), rng.uniform(5, 20))
             for i in range(n_lines)]
    t0 = time.time()
    kept = lstconv.delete_bad_score(lines, 10.0)
    t_f = time.time() - t0
    t0 = time.time()
    star = lstconv.lst_to_star(lines, 1, 1.0, "/data/")
    t_s = time.time() - t0
    out["lst"] = {"lines": n_lines, "score_filter_s": t_f, "kept": len(kept),
                  "lst_to_star_s": t_s, "star_bytes": len(star)}
    log(f"[bench] io: score filter of {n_lines} lines {t_f * 1e3:.0f} ms ({len(kept)} kept), "
        f"lst -> STAR {t_s * 1e3:.0f} ms")
    return out


def dedupe_rate(picks, n_views: int, no_cpu: bool, center: float = 20.0,
                euler_thres: float = 10.0):
    """Duplicate removal (gs_remove_repeats) of the run's pick list, one
    micrograph: Euler angles of the synthetic views, the printed psi as phi.
    The oracle's single-thread time on the same list is reported beside it."""
    import gisspa_amd
    from gisspa_amd import synth
    eul = synth.sphere_eulers(n_views)
    v = picks["view"]
    recs = gisspa_amd.pick_records(np.zeros(len(picks), np.int32), picks["cx"], picks["cy"],
                                   eul[v, 0], eul[v, 1], picks["psi"], picks["score"])
    gisspa_amd.remove_repeats(recs[:64], center, euler_thres)      # warm-up
    t = time.time()
    keep = gisspa_amd.remove_repeats(recs, center, euler_thres)
    gpu_ms = (time.time() - t) * 1e3
    out = {"picks": int(len(recs)), "kept": int(keep.sum()), "center": center,
           "euler_thres": euler_thres, "gpu_ms": gpu_ms, "cpu_ms": None, "identical": None}
    if not no_cpu:
        from oracle import oracle as orc
        t = time.time()
        ref = orc.remove_repeats(recs["group"], recs["x"], recs["y"], recs["alt"], recs["az"],
                                 recs["phi"], recs["score"], center, euler_thres)
        out["cpu_ms"] = (time.time() - t) * 1e3
        out["identical"] = bool(np.array_equal(ref, keep))
    return out


def project_rate(n_views: int, no_cpu: bool, box: int = 180):
    """Template projection (gs_project) of the full workload's 180^3 density
    (the 60-blob map its templates come from) at all n_views views; the
    oracle on 2 of them, extrapolated."""
    import gisspa_amd
    from gisspa_amd import synth
    vol = synth.render_volume(synth.blob_density(60, box, 1), box)
    eul = np.zeros((n_views, 3), np.float32)
    eul[:, :2] = synth.sphere_eulers(n_views)
    gisspa_amd.project(vol, eul[:8])                       # warm-up
    t = time.time()
    gisspa_amd.project(vol, eul)
    gpu_s = time.time() - t
    out = {"views": n_views, "box": box, "gpu_ms": gpu_s * 1e3, "views_per_s": n_views / gpu_s,
           "samples_per_s": n_views * box ** 3 / gpu_s, "cpu_views_per_s": None}
    if not no_cpu:
        from oracle import oracle as orc
        t = time.time()
        orc.project(vol, eul[:2].astype(np.float64))
        out["cpu_views_per_s"] = 2 / (time.time() - t)
    return out


BATCH_SIZES = [(3838, 3710), (4096, 4096), (5760, 4092)]


def _batched_micrograph(args):
    i, seed, nx, ny, df, npart = args
    from gisspa_amd import synth
    dens = synth.blob_density(60, 180, 1)
    mg = synth.make_micrograph(nx, ny, dens, reference_views(), npart, 0.05, seed * 100 + i,
                               apix=1.0, defocus_um=df, ampconst=0.1, box=180, psi_step=5.0)
    return mg.image


def batched_inputs(n_mics: int, seed: int = 1):
    """n_mics micrographs of sizes drawn from 3838x3710, 4096x4096 and
    5760x4092, each with its own defocus (0.8-2.5 um) and 200-400 particles
    at the reference's 1649 views; made in worker processes (spawned, so
    they never touch the GPU) before the device is used."""
    import multiprocessing as mp
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_mics):
        nx, ny = BATCH_SIZES[int(rng.integers(0, 3))]
        jobs.append((i, seed, nx, ny, float(rng.uniform(0.8, 2.5)), int(rng.integers(200, 401))))
    t = time.time()
    with mp.get_context("spawn").Pool(min(16, os.cpu_count() or 1, n_mics)) as pool:
        imgs = pool.map(_batched_micrograph, jobs)
    log(f"[bench] batched: generated {n_mics} micrographs in {time.time() - t:.1f}s")
    return imgs, [j[4] for j in jobs], [f"{j[2]}x{j[3]}" for j in jobs]


def batched_rate(templates, inputs, npsi: int, refine_per_mic: int, refine_iters: int,
                 tile_block: int = 0, autotune: bool = False):
    """Many micrographs in one tile pool (gs_set_micrographs): the batch of
    batched_inputs, templates of the 1649 views in mixed boxes of 128, 180 and
    256 px (cropped / padded from the 180 px projections).  Timed after one
    warm-up psi: npsi in-plane steps, then the local refinement of the
    top refine_per_mic picks per micrograph of those steps (perturbed by
    +-2.5 deg, +-1 px, +-200 A).  Rates: the picker alone (the steps
    extrapolated to 72) and picking + refinement per batch."""
    import gisspa_amd
    imgs, dfs, shapes = inputs
    n_mics = len(imgs)
    mixed = []
    for j, tm in enumerate(templates):
        b = (128, 180, 256)[j % 3]
        if b <= 180:
            o = (180 - b) // 2
            mixed.append(np.ascontiguousarray(tm[o:o + b, o:o + b]))
        else:
            z = np.zeros((b, b), np.float32)
            o = (b - 180) // 2
            z[o:o + 180, o:o + 180] = tm
            mixed.append(z)
    pk = gisspa_amd.Picker(0, **FULL_PARAMS)
    if tile_block:
        pk.set_tile_block(tile_block)
    pk.set_templates_mixed(mixed, dfs[0])
    t = time.time()
    pk.set_micrographs(imgs, dfs)
    t_prep = time.time() - t
    info = pk.info()
    tune = pk.autotune(0) if autotune else None
    pk.pick(0, 1)                      # warm-up
    rng = np.random.default_rng(11)
    t = time.time()
    picks = pk.pick(1, 1 + npsi)
    t_pick = time.time() - t
    # the best picks of each micrograph, at the picked grid parameters,
    # perturbed as the differentiation workload's candidates
    sel = np.concatenate([np.flatnonzero(picks["mic"] == m)[
        np.argsort(-picks["score"][picks["mic"] == m])[:refine_per_mic]] for m in range(n_mics)])
    top = picks[np.sort(sel)]
    peaks = gisspa_amd.peaks_from_picks(top, FULL_PARAMS["phistep"], 0.0)
    peaks["defocus"] = np.asarray(dfs, np.float32)[top["mic"]]
    peaks["psi"] += rng.uniform(-2.5, 2.5, len(peaks)).astype(np.float32)
    peaks["dx"] += rng.uniform(-1, 1, len(peaks)).astype(np.float32)
    peaks["dy"] += rng.uniform(-1, 1, len(peaks)).astype(np.float32)
    peaks["defocus"] += rng.uniform(-0.02, 0.02, len(peaks)).astype(np.float32)
    _, fin = pk.refine(peaks, refine_iters)
    t_all = time.time() - t
    t_refine = t_all - t_pick
    pk.close()
    t_psi = t_pick / npsi
    sec = t_prep + 72 * t_psi
    sec_refined = sec + t_refine
    full_equiv = info.n_tiles / n_mics / 88.0   # mean tiles per micrograph over 88
    return {"micrographs": n_mics, "sizes": shapes, "tiles": info.n_tiles,
            "maps_per_psi": info.n_tiles * templates.shape[0], "prep_s": t_prep,
            "psi_s": t_psi, "picks_timed": int(len(picks)),
            "value": n_mics * 3600.0 / sec,
            "refined_value": n_mics * 3600.0 / sec_refined,
            "refine_candidates": int(len(peaks)), "refine_s": t_refine,
            "mean_score_refined": float(fin["score"].mean()),
            "full_size_equivalent_value": n_mics * 3600.0 / sec * full_equiv,
            "full_size_equivalent_refined_value": n_mics * 3600.0 / sec_refined * full_equiv,
            "tile_block": tune.tile_block if tune else tile_block,
            "unit": "micrographs/hour",
            "sample": f"{npsi} psi timed after one warm-up psi, extrapolated to 72; refinement "
                      f"of the top {refine_per_mic} picks per micrograph ({len(peaks)} "
                      f"candidates, up to {refine_iters} steps) timed with the picking; "
                      f"full-size equivalent = rate x mean tiles per micrograph / 88"}


def rocfft_compare(L: int, batch: int = 189, reps: int = 5):
    """Vendor FFT timing comparison (torch.fft on ROCm runs rocFFT): one
    batched inverse 2D transform per map, as C2R from the half spectrum (the
    transform inside the fused correlation) and as full complex C2C (the
    reference's cuFFT call, main.cu:848).  Only the transform: no product,
    no peak search, no partial sums."""
    import torch
    dev = torch.device("cuda:0")
    H = L // 2 + 1
    out = {"maps_per_call": batch}
    half = torch.randn(batch, L, H, dtype=torch.complex64, device=dev)
    full = torch.randn(batch, L, L, dtype=torch.complex64, device=dev)
    for name, fn, nbytes in (("c2r", lambda: torch.fft.irfft2(half, s=(L, L)),
                              8 * L * H + 4 * L * L),
                             ("c2c", lambda: torch.fft.ifft2(full), 16 * L * L)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps / batch
        out[f"{name}_us_per_map"] = us
        out[f"{name}_gbs"] = nbytes / (us * 1e-6) / 1e9
    del half, full
    torch.cuda.empty_cache()
    return out


def stress_rate(n_mics: int, npsi: int, n_refine: int, refine_iters: int, seed: int = 3):
    """The combined stress shape: n_mics 5760x4092 micrographs (defocus
    0.8-3.0 um, 300 particles each) in one tile pool, 256 px templates of a
    60-blob 256^3 density at 4600 views of a ~3 degree sphere made by the GPU
    projector (gs_project), 3 degree in-plane steps (120 psi); npsi steps
    timed (after one warm-up step) and extrapolated to 120, then local
    refinement of the top n_refine picks of the timed steps."""
    import gisspa_amd
    from gisspa_amd import synth
    rng = np.random.default_rng(seed)
    box, nviews = 256, 4600
    t = time.time()
    dens = synth.blob_density(60, box, 1)
    vol = synth.render_volume(dens, box)
    views = synth.sphere_eulers(nviews)
    eul = np.zeros((nviews, 3), np.float32)
    eul[:, :2] = views
    t1 = time.time()
    templates = gisspa_amd.project(vol, eul)
    t_proj = time.time() - t1
    log(f"[bench] stress: {nviews} templates of {box}^2 projected in {t_proj:.2f}s")
    imgs, dfs = [], []
    for i in range(n_mics):
        df = float(rng.uniform(0.8, 3.0))
        mg = synth.make_micrograph(5760, 4092, dens, views, 300, 0.05, seed * 100 + i, apix=1.0,
                                   defocus_um=df, ampconst=0.1, box=box, psi_step=3.0)
        imgs.append(mg.image)
        dfs.append(df)
        log(f"[bench] stress: micrograph {i} defocus {df:.2f} um")
    log(f"[bench] stress: inputs in {time.time() - t:.1f}s")
    # a lower threshold than the headline run, so that two timed steps hold
    # the 50,000 peaks the refinement takes
    params = dict(FULL_PARAMS, phistep=3.0, diameter=0.75 * box, threshold=8.0)
    pk = gisspa_amd.Picker(0, **params)
    pk.set_templates(templates, dfs[0])
    t = time.time()
    pk.set_micrographs(imgs, dfs)
    t_prep = time.time() - t
    info = pk.info()
    pk.pick(0, 1)                      # warm-up
    t = time.time()
    picks = pk.pick(1, 1 + npsi)
    t_psi = (time.time() - t) / npsi
    out = {"micrographs": n_mics, "views": nviews, "box": box, "psi": info.n_psi,
           "tiles": info.n_tiles, "maps_per_psi": info.n_tiles * nviews, "prep_s": t_prep,
           "psi_s": t_psi, "project_s": t_proj,
           "value": n_mics * 3600.0 / (t_prep + info.n_psi * t_psi), "unit": "micrographs/hour",
           "sample": f"{npsi} of {info.n_psi} psi timed, extrapolated", "picks_timed": len(picks)}
    if n_refine and len(picks):
        top = picks[np.argsort(-picks["score"])[:n_refine]]
        peaks = gisspa_amd.peaks_from_picks(top, params["phistep"], 0.0)
        peaks["defocus"] = np.asarray(dfs, np.float32)[top["mic"]]
        pert = peaks.copy()
        pert["psi"] += rng.uniform(-2.5, 2.5, len(pert)).astype(np.float32)
        pert["dx"] += rng.uniform(-1, 1, len(pert)).astype(np.float32)
        pert["dy"] += rng.uniform(-1, 1, len(pert)).astype(np.float32)
        pert["defocus"] += rng.uniform(-0.02, 0.02, len(pert)).astype(np.float32)
        s0 = pk.adjoint(pert)["score"]
        t = time.time()
        _, fin = pk.refine(pert, refine_iters)
        tr = time.time() - t
        out["refine"] = {"candidates": int(len(pert)), "seconds": tr,
                         "evaluations": int(pk.timing().refine_evals),
                         "mean_score_perturbed": float(s0.mean()),
                         "mean_score_refined": float(fin["score"].mean())}
    pk.close()
    return out


def cli_rate(templates, image, defocus, n_mics: int, cache_dir: str, groups=(1, 4)):
    """The ./main command line end to end on n_mics full-size micrographs
    (the benchmark micrograph shifted by a different offset for each, defocus
    spread 1.6-2.0 um), files in the reference's formats; one run per
    micrograph_batch value in groups.  Reports the program's own loop time
    (main.cu:1271-1273 prints the same span) and the process wall time
    (template file reading and preparation included)."""
    import re
    import subprocess

    import gisspa_amd
    from gisspa_amd import synth
    d = os.path.join(cache_dir, "gisspa_cli_bench")
    mics = []
    for i in range(n_mics):
        mics.append((np.roll(image, (97 * i, 211 * i), axis=(0, 1)),
                     float(defocus + 0.4 * (i / max(1, n_mics - 1) - 0.5))))
    wl = synth.Workload("cli", templates, synth.sphere_eulers(templates.shape[0]), mics,
                        dict(FULL_PARAMS), [])
    t = time.time()
    out = {"micrographs": n_mics, "runs": {}}
    outputs = {}
    for g in groups:
        cfg = synth.write_workload(wl, d, extra=dict(micrograph_batch=g))
        if g == groups[0]:
            log(f"[bench] cli: wrote {n_mics} micrographs and the templates in "
                f"{time.time() - t:.1f}s")
        t0 = time.time()
        r = subprocess.run([gisspa_amd.MAIN_PATH, cfg], capture_output=True, text=True,
                           timeout=1200)
        wall = time.time() - t0
        if r.returncode != 0:
            raise RuntimeError(f"./main failed ({r.returncode}): {r.stderr[-2000:]}")
        m = re.search(r"Total consumed time is: ([\d.]+) seconds", r.stdout)
        loop = float(m.group(1))
        with open(os.path.join(d, "picks.lst")) as f:
            outputs[g] = f.read()
        out["runs"][str(g)] = {"loop_s": loop, "wall_s": wall,
                               "micrographs_per_hour": n_mics * 3600.0 / loop,
                               "lines": outputs[g].count("\n")}
        log(f"[bench] cli: micrograph_batch {g}: loop {loop:.1f}s, wall {wall:.1f}s, "
            f"{n_mics * 3600.0 / loop:.1f} micrographs/hour")
    out["identical_outputs"] = len(set(outputs.values())) == 1
    out["unit"] = "micrographs/hour"
    out["sample"] = f"{n_mics} micrographs, all 72 psi, one run per micrograph_batch"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--psi-limit", type=int, default=0,
                    help="exploration/profiling only: psi steps per micrograph (0 = all 72)")
    ap.add_argument("--batch", type=int, default=0, help="CC maps per launch (0 = library default)")
    ap.add_argument("--graphs", action="store_true", help="replay per-psi HIP graphs")
    ap.add_argument("--tile-block", type=int, default=0, help="tiles per block (0 = default)")
    ap.add_argument("--pipeline", type=int, default=3,
                    help="0: one launch per psi, 1: one launch per batch, 2: two streams, "
                         "3: pass-1 / pass-2 launches on one stream")
    ap.add_argument("--lag", type=int, default=0)
    ap.add_argument("--ring", type=int, default=0)
    ap.add_argument("--threshold", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-psi", type=int, default=1)
    ap.add_argument("--cpu-views", type=int, default=8)
    ap.add_argument("--refine", type=int, default=10000,
                    help="refine the top N picks (perturbed by +-2.5 deg, +-1 px, +-200 A) "
                         "for --refine-iters steps after the timed run (0 = skip)")
    ap.add_argument("--refine-iters", type=int, default=20)
    ap.add_argument("--dedupe", type=int, default=1,
                    help="time the duplicate removal of the pick list (0 = skip)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="choose the correlation schedule on the device before timing (0 = the "
                         "built-in defaults)")
    ap.add_argument("--stress", type=int, default=0, metavar="N",
                    help="also run the combined stress shape on N 5760x4092 micrographs "
                         "(4600 views of 256 px, 120 psi, refinement of the top 50,000 picks; "
                         "0 = skip)")
    ap.add_argument("--batched", type=int, default=16, metavar="N",
                    help="also time N micrographs of uneven sizes in one tile pool with mixed "
                         "template boxes, picking and refinement (0 = skip)")
    ap.add_argument("--batched-psi", type=int, default=4,
                    help="in-plane steps timed in the batched leg (after one warm-up step)")
    ap.add_argument("--batched-refine", type=int, default=500, metavar="N",
                    help="picks refined per micrograph in the batched leg")
    ap.add_argument("--cli", type=int, default=0, metavar="N",
                    help="also run the ./main command line on N full-size micrographs, once per "
                         "micrograph_batch of 1 and 4 (0 = skip)")
    ap.add_argument("--norm1", type=int, default=206, metavar="VIEWS",
                    help="also time the norm_type 1 picker with the first VIEWS views, scaled "
                         "to all (0 = skip)")
    ap.add_argument("--io", type=int, default=0,
                    help="also time the host-side formats: MRC / HDF5 reading, score filter, "
                         "lst -> STAR (0 = skip)")
    ap.add_argument("--small", type=int, default=1,
                    help="also run the small configuration (1024^2, 64 views, 36 psi) in full "
                         "on the GPU and, unless --no-cpu, on the FP64 oracle (0 = skip)")
    ap.add_argument("--rocfft", type=int, default=1,
                    help="time rocFFT's inverse 2D transforms of the same size (0 = skip)")
    ap.add_argument("--project", type=int, default=1,
                    help="time the template projection of the density (0 = skip)")
    ap.add_argument("--time-kernels", type=int, default=0, metavar="REPS",
                    help="only time each correlation pass alone (gs_time_kernels, REPS launches "
                         "of each, one at a time on one stream) for counter runs, then exit "
                         "without a JSON line (0 = the normal run)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--cache", default=os.environ.get("TMPDIR", "/tmp"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the picker on one GPU; --gpus {args.gpus} is not "
              f"supported (shard micrographs across processes instead)", file=sys.stderr)
        sys.exit(2)

    # the batched leg's micrographs, made before the device is used
    batched_in = batched_inputs(args.batched) if args.batched else None

    import gisspa_amd
    FULL_PARAMS["threshold"] = args.threshold
    templates, image, defocus = full_workload(args.cache)
    ny, nx = image.shape
    hip = Hip()
    d_img = hip.malloc(image.nbytes)
    hip.h2d(d_img, np.ascontiguousarray(image, dtype=np.float32))
    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        a, b = hip.malloc(nb), hip.malloc(nb)
        hip.d2d(b, a, nb)
        hip.sync()
        hip.free(a)
        hip.free(b)
        log(f"[bench] calibration copy of {nb} bytes done")

    pk = gisspa_amd.Picker(0, **FULL_PARAMS)
    if args.batch:
        pk.set_batch(args.batch)
    if args.tile_block:
        pk.set_tile_block(args.tile_block)
    pk.set_pipeline(args.pipeline, args.lag, args.ring)
    pk.set_graphs(args.graphs)
    pk.set_profiling(True)
    t = time.time()
    pk.set_templates(templates, defocus)
    log(f"[bench] template prep {time.time() - t:.2f}s")
    if args.time_kernels:
        pk.set_micrograph_device(d_img, nx, ny, defocus)
        k1, k2, nb = pk.time_kernels(0, args.time_kernels)
        log(f"[bench] time_kernels: {nb} maps per launch: pass 1 {k1:.3f} ms "
            f"({k1 / nb * 1e3:.3f} us/map), pass 2 {k2:.3f} ms ({k2 / nb * 1e3:.3f} us/map)")
        pk.close()
        hip.free(d_img)
        return
    npsi = 72 if args.psi_limit <= 0 else min(72, args.psi_limit)
    tune = None
    pipeline = args.pipeline
    if args.autotune:
        t = time.time()
        pk.set_micrograph_device(d_img, nx, ny, defocus)
        tr = pk.autotune(0)
        tune = {"pass1_variant": tr.pass1_variant, "pass2_variant": tr.pass2_variant,
                "batch": tr.batch, "pipeline": tr.pipeline, "psi_ms": tr.psi_ms,
                "pass1_us_per_map": list(tr.pass1_us), "pass2_us_per_map": list(tr.pass2_us),
                "seconds": time.time() - t}
        log(f"[bench] autotune ({tune['seconds']:.1f}s): FFT variants {tr.pass1_variant}/"
            f"{tr.pass2_variant}, batch {tr.batch}, pipeline {tr.pipeline}, psi {tr.psi_ms:.1f} ms;"
            f" pass 1 us/map {[round(v, 3) for v in tr.pass1_us]}, pass 2 "
            f"{[round(v, 3) for v in tr.pass2_us]}")
        pipeline = tr.pipeline

    def step():
        pk.set_micrograph_device(d_img, nx, ny, defocus)
        return pk.pick(0, npsi)

    for i in range(args.warmup):
        t = time.time()
        picks = step()
        log(f"[bench] warmup {i}: {time.time() - t:.2f}s, {len(picks)} picks")
    hip.sync()
    t0 = time.time()
    tms = tms2 = tpsi = 0.0
    nl = nmaps = 0
    for i in range(args.steps):
        picks = step()
        tm = pk.timing()
        tms += tm.corr_ms
        tms2 += tm.corr_pass2_ms
        tpsi += tm.corr_psi_ms
        nl += tm.corr_launches
        nmaps += tm.corr_sampled_maps
        log(f"[bench] step {i}: {time.time() - t0:.2f}s cumulative, {len(picks)} picks, "
            f"prep {tm.micrograph_prep_ms:.1f} ms, pick {tm.pick_total_ms:.1f} ms")
    hip.sync()
    dt = time.time() - t0
    sec_per_step = dt / args.steps
    sec_per_mic = sec_per_step * (72 / npsi)
    mph = 3600.0 / sec_per_mic
    info = pk.info()
    T = info.n_tiles
    N = info.n_views
    L = info.l
    H = L // 2 + 1
    refine = None
    if args.refine and len(picks):
        top = picks[np.argsort(-picks["score"])[:args.refine]]
        base = gisspa_amd.peaks_from_picks(top, FULL_PARAMS["phistep"], defocus)
        rng = np.random.default_rng(0)
        pert = base.copy()
        pert["psi"] += rng.uniform(-2.5, 2.5, len(pert)).astype(np.float32)
        pert["dx"] += rng.uniform(-1, 1, len(pert)).astype(np.float32)
        pert["dy"] += rng.uniform(-1, 1, len(pert)).astype(np.float32)
        pert["defocus"] += rng.uniform(-0.02, 0.02, len(pert)).astype(np.float32)
        s_grid = pk.adjoint(base)["score"]
        s0 = pk.adjoint(pert)["score"]
        hip.sync()
        t = time.time()
        _, fin = pk.refine(pert, args.refine_iters)
        hip.sync()
        tr = time.time() - t
        nev = pk.timing().refine_evals
        refine = {"candidates": int(len(pert)), "iters": args.refine_iters, "seconds": tr,
                  "candidates_per_s": len(pert) / tr, "evaluations": int(nev),
                  "evaluations_per_s": nev / tr,
                  "mean_score_perturbed": float(s0.mean()),
                  "mean_score_refined": float(fin["score"].mean()),
                  "mean_score_at_picks": float(s_grid.mean())}
        log(f"[bench] refine: {len(pert)} candidates, up to {args.refine_iters} steps "
            f"({nev} evaluations) in {tr:.2f}s; "
            f"mean objective {s0.mean():.3f} -> {fin['score'].mean():.3f} "
            f"(at the picks {s_grid.mean():.3f})")
    pk.close()

    dedupe = None
    if args.dedupe and len(picks):
        dedupe = dedupe_rate(picks, N, args.no_cpu)
        log(f"[bench] dedupe: {dedupe['picks']} picks -> {dedupe['kept']} in "
            f"{dedupe['gpu_ms']:.2f} ms" + (f" (oracle {dedupe['cpu_ms']:.0f} ms, 1 thread)"
                                              if dedupe.get("cpu_ms") else ""))

    proj = None
    if args.project:
        proj = project_rate(N, args.no_cpu)
        log(f"[bench] project: {proj['views']} views of {proj['box']}^3 in {proj['gpu_ms']:.1f} ms"
            + (f" (oracle {proj['cpu_views_per_s']:.2f} views/s, 1 thread)"
               if proj["cpu_views_per_s"] else ""))

    # Roofline of the correlation kernel, from hipEvents bracketing every
    # correlation launch of the first psi of each timed step on the stream it
    # runs on.  Algorithmic bytes per CC map: 24 L H = pass 1 (template half
    # spectrum read 8 L H + intermediate written 8 L H) + pass 2 (intermediate
    # read 8 L H).
    mpl = nmaps / max(nl, 1)                   # CC maps per launch
    if pipeline in (2, 3):                     # separate pass-1 / pass-2 launches
        t1, t2 = tms / max(nl, 1), tms2 / max(nl, 1)
        a1 = 16.0 * L * H * mpl / (t1 * 1e-3) / 1e9
        a2 = 8.0 * L * H * mpl / (t2 * 1e-3) / 1e9
        kern, t_launch, ach, bpm = (("k_corr_pass1", t1, a1, 16.0 * L * H) if t1 >= t2 else
                                    ("k_corr_pass2", t2, a2, 8.0 * L * H))
        log(f"[bench] k_corr_pass1 {t1:.4f} ms/launch ({a1:.0f} GB/s), k_corr_pass2 {t2:.4f} "
            f"ms/launch ({a2:.0f} GB/s), {mpl:.1f} maps/launch"
            + (" (two streams overlap)" if pipeline == 2 else " (one stream)"))
    else:
        kern, t_launch, bpm = ("k_corr_stream" if pipeline == 0 else "k_corr_fused",
                               tms / max(nl, 1), 24.0 * L * H)
        ach = bpm * mpl / (t_launch * 1e-3) / 1e9
        log(f"[bench] {kern} {t_launch:.4f} ms/launch, {mpl:.1f} maps/launch, "
            f"{ach:.0f} GB/s algorithmic, {nl} launches sampled")
    # both passes of the sampled psi against its correlation wall time
    pair = (24.0 * L * H * nmaps / (tpsi * 1e-3) / 1e9) if tpsi > 0 else None
    if pair:
        log(f"[bench] psi correlation: {tpsi / max(args.steps, 1):.1f} ms, {pair:.0f} GB/s "
            f"algorithmic ({pair / HBM_PEAK_GBS:.3f} of HBM peak)")

    batched = None
    if args.batched:
        batched = batched_rate(templates, batched_in, max(1, args.batched_psi),
                               args.batched_refine, args.refine_iters,
                               tile_block=args.tile_block, autotune=bool(args.autotune))
        batched_in = None
        log(f"[bench] batched: {batched['micrographs']} micrographs ({batched['tiles']} tiles), "
            f"{batched['psi_s']:.2f}s per psi, {batched['value']:.1f} micrographs/hour picked, "
            f"{batched['refined_value']:.1f} picked + refined ({batched['refine_candidates']} "
            f"candidates in {batched['refine_s']:.2f}s); full-size equivalent "
            f"{batched['full_size_equivalent_value']:.1f} / "
            f"{batched['full_size_equivalent_refined_value']:.1f}")

    rocfft = None
    if args.rocfft:
        rocfft = rocfft_compare(L)
        if pair:
            rocfft["fused_correlation_us_per_map"] = tpsi / max(args.steps, 1) * 1e3 / (
                nmaps / max(args.steps, 1))
        log(f"[bench] rocFFT inverse 2D {L}^2: C2R {rocfft['c2r_us_per_map']:.3f} us/map "
            f"({rocfft['c2r_gbs']:.0f} GB/s), C2C {rocfft['c2c_us_per_map']:.3f} us/map; fused "
            f"correlation (products, transforms, peak search) "
            f"{rocfft.get('fused_correlation_us_per_map', 0):.3f} us/map")

    stress = None
    if args.stress:
        stress = stress_rate(args.stress, max(1, min(npsi, 2)), 50000, args.refine_iters)
        log(f"[bench] stress: {stress['micrographs']} micrographs, {stress['maps_per_psi']} maps "
            f"per psi, {stress['psi_s']:.2f}s per psi, {stress['value']:.1f} micrographs/hour"
            + (f"; refine {stress['refine']['candidates']} in {stress['refine']['seconds']:.1f}s"
               if stress.get("refine") else ""))

    io = io_rate(templates, image, args.cache) if args.io else None

    cli = None
    if args.cli:
        cli = cli_rate(templates, image, defocus, args.cli, args.cache)

    copy_gbs = hip.copy_rate(4 << 30) / 1e9
    log(f"[bench] 4 GiB device copy: {copy_gbs:.0f} GB/s (read + write); psi correlation at "
        f"{(pair or 0) / copy_gbs:.3f} of it")

    plain = None
    if not args.no_plain:
        plain = plain_rate(templates, d_img, nx, ny, defocus, args.plain_psi)
        log(f"[bench] plain: {plain['sample']}: {plain['value']:.2f} micrographs/hour")
    prep = prep_times(templates, d_img, nx, ny, defocus, hip)
    log(f"[bench] micrograph prep: phase_flip 0 {prep['phase_flip_0_ms']:.1f} ms, phase_flip 1 "
        f"{prep['phase_flip_1_ms']:.1f} ms")

    norm1 = None
    if args.norm1:
        norm1 = norm1_rate(templates, d_img, nx, ny, defocus, args.norm1)
        log(f"[bench] norm_type 1: {norm1['sample']}: {norm1['value']:.2f} micrographs/hour")
    hip.free(d_img)

    small = None
    if args.small:
        small, small_wl, small_picks = small_config()
        log(f"[bench] small configuration ({small['maps']} maps): GPU {small['gpu_s']:.3f} s "
            f"({small['picks']} picks)")

    cpu = None
    if not args.no_cpu:
        nthreads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(templates, image, defocus, nthreads, args.cpu_views)
        log(f"[bench] cpu baseline: {cpu['sample']}")
        if small is not None:
            small["cpu_s"], small["identical_picks"] = cpu_baseline_small(small_wl, small_picks,
                                                                         nthreads)
            small["cpu_threads"] = nthreads
            log(f"[bench] small configuration: FP64 oracle {small['cpu_s']:.1f} s on {nthreads} "
                f"threads, identical picks {small['identical_picks']}")

    out = {
        "metric": "micrographs/hour at 5760x4092, 1650 views x72 psi; batched vs plain",
        "value": mph,
        "unit": "micrographs/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "full-size: one 5760x4092 micrograph, 1649 views x 72 psi, "
                               "180px templates, window 720 overlap 180 (88 tiles)",
                   "views": N, "psi": 72, "tiles": T, "window": L, "psi_timed": npsi,
                   "maps_per_micrograph": N * 72 * T, "threshold": args.threshold,
                   "mode": "fused" + ("+graphs" if args.graphs else ""), "autotune": tune},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC_PER_MAP_720[kern] * mpl)
                     if (L == 720 and TRAFFIC_PER_MAP_720 and kern in TRAFFIC_PER_MAP_720)
                     else None,
                     "launch_ms": t_launch, "maps_per_launch": mpl, "bytes_per_map": bpm,
                     "psi_correlation_achieved": pair,
                     "psi_correlation_frac": (pair / HBM_PEAK_GBS) if pair else None,
                     "copy_gbs": copy_gbs,
                     # each pass alone (autotune's timing of the chosen FFT schedule, one
                     # stream, no overlap): algorithmic bytes / isolated time per map
                     "pass1_isolated_gbs": (16.0 * L * H / (tune["pass1_us_per_map"]
                                            [tune["pass1_variant"]] * 1e-6) / 1e9)
                     if tune else None,
                     "pass2_isolated_gbs": (8.0 * L * H / (tune["pass2_us_per_map"]
                                            [tune["pass2_variant"]] * 1e-6) / 1e9)
                     if tune else None},
        "refine": refine,
        "dedupe": dedupe,
        "batched": batched,
        "stress": stress,
        "cli": cli,
        "norm1": norm1,
        "prep": prep,
        "small": small,
        "io": io,
        "rocfft": rocfft,
        "project": proj,
        "plain": plain,
        "speedup_vs_plain": (mph / plain["value"]) if plain else None,
        # picking plus the local refinement of the top picks (the refine leg's
        # candidates) per micrograph
        "refined_value": (3600.0 / (sec_per_mic + refine["seconds"])) if refine else None,
        "refined_speedup_vs_plain": (3600.0 / (sec_per_mic + refine["seconds"]) / plain["value"])
        if (refine and plain) else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
