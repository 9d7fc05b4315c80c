"""GPU parity at the shapes of the batched, differentiation and combined
stress workloads, on slices small enough for the FP64 oracle.

  * Batched: one tile pool of a 3838 x 3710, a 4096 x 4096 and a
    5760 x 4092 micrograph (201 tiles at window 720), each with its own
    defocus, templates of mixed boxes 128 / 180 / 256 at views of the
    reference's euler_C1_step5.lst; per micrograph, every per-map score
    within 1e-4 * max(|s|, 1) of the oracle's, identical peak positions and
    pick list (the reference searches one micrograph per loop iteration,
    main.cu:1149-1269, and centres each template by its own box,
    main.cu:182-223).
  * Differentiation: candidates from the full-size micrograph's CC maps
    (5760 x 4092, window 720, all 72 in-plane steps), perturbed by up to
    +-2.5 degrees, +-1 px and +-200 A; objective within 1e-4 of the oracle's
    and gradient within 1e-3 of FP64 central differences on 128 of them
    (steps as in test_gpu_refine.py: the objective has kinks in psi and
    defocus that a 1e-3 step would straddle); after 20 refinement steps the
    objective has not dropped and equals the oracle's at the final
    parameters.
  * Combined stress: 256 px templates, 3 degree in-plane steps (120), eight
    views of a ~3 degree sphere on a 5760 x 4092 micrograph, two in-plane
    steps against the oracle, then 64 refinements checked as above (the
    gradient on 16 of them).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gisspa_amd import synth
from oracle import oracle as orc
from tests import _cases
from tests.test_gpu_refine import FD_STEPS

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
PARAMS = dict(angpix=1.0, phistep=5.0, kk=3, energy=300, cs=2.7, Highres=8, Lowres=100,
              diameter=135.0, threshold=7.0, window_size=720, overlap=180, ampconst=0.1)


def _compare_maps(got, ref, what):
    rel = np.abs(got["score"].astype(np.float64) - ref["score"]) / np.maximum(
        np.abs(ref["score"]), 1.0)
    mism = np.count_nonzero(got["pos"] != ref["pos"])
    print(f"{what}: {ref.size} maps, score max rel err {rel.max():.2e}, position mismatches "
          f"{mism}", flush=True)
    assert rel.max() < 1e-4
    assert mism == 0


def _compare_picks(picks, sel, what):
    print(f"{what}: picks gpu {len(picks)} oracle {len(sel)}", flush=True)
    assert len(picks) == len(sel)
    for k in ("psi_index", "view"):
        assert np.array_equal(picks[k], sel[k])
    assert np.array_equal(picks["cx"], sel["cx"].astype(np.float32))
    assert np.array_equal(picks["cy"], sel["cy"].astype(np.float32))


def test_batched_pool_of_uneven_micrographs(gpu):
    import gisspa_amd
    views = _cases.c1_step5_views()
    vsel = views[[0, 400, 800, 1200]]
    dens = synth.blob_density(60, 180, 1)
    boxes = [128, 180, 256, 180]
    tms = [synth.project_templates(dens, b, vsel[j:j + 1])[0] for j, b in enumerate(boxes)]
    sizes = [(3838, 3710), (4096, 4096), (5760, 4092)]
    dfs = [0.9, 1.7, 2.4]
    imgs = [synth.make_micrograph(nx, ny, dens, vsel, 200 + 100 * i, 0.05, 50 + i, apix=1.0,
                                  defocus_um=df, ampconst=0.1, box=180, psi_step=5.0).image
            for i, ((nx, ny), df) in enumerate(zip(sizes, dfs))]
    with gisspa_amd.Picker(0, **PARAMS) as pk:
        pk.set_templates_mixed(tms, dfs[0])
        pk.set_micrographs(imgs, dfs)
        info = pk.info()
        picks = pk.pick(0, 1)
        st = pk.mapstats().reshape(1, info.n_tiles, len(tms))
    assert info.n_mics == 3 and info.n_tiles == 49 + 64 + 88 and info.l == 720
    L = 720
    padded = np.zeros((len(tms), L, L), np.float32)
    for j, t in enumerate(tms):
        b = t.shape[0]
        o = (L - b) // 2
        padded[j, o:o + b, o:o + b] = t
    op = orc.params(L, **PARAMS)
    t0 = 0
    for i, (img, (nx, ny)) in enumerate(zip(imgs, sizes)):
        # phase_flip 0: the micrograph's own CTF is not used; the templates
        # carry the first micrograph's (main.cu:1104-1133)
        ref = orc.pick_norm0(op, orc.ctf(dfs[0]), padded, img, 0, 1, nthreads=NTHREADS)
        nt = ref.shape[1]
        _compare_maps(st[:, t0:t0 + nt, :], ref, f"micrograph {i} ({nx}x{ny})")
        t0 += nt
        mine = picks[picks["mic"] == i]
        _compare_picks(mine, orc.select_norm0(op, ref, 0, nx, ny), f"micrograph {i}")


def _perturb(base, seed, ddf=0.02):
    rng = np.random.default_rng(seed)
    p = base.copy()
    p["psi"] += rng.uniform(-2.5, 2.5, len(p)).astype(np.float32)
    p["dx"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
    p["dy"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
    p["defocus"] += rng.uniform(-ddf, ddf, len(p)).astype(np.float32)
    return p


def _oracle_objective(op, df, white, tiles, offs, peaks, fd_rows=()):
    """Oracle objective of every candidate and FP64 central differences for
    the rows in fd_rows (threads over candidates: or_refine_score is
    re-entrant)."""
    ox, oy = offs

    def one(i):
        q = peaks[i]
        t = int(q["tile"])
        args = (op, orc.ctf(df), white[int(q["view"])], tiles[t],
                float(q["cx"]) - ox[t % len(ox)], float(q["cy"]) - oy[t // len(ox)])
        theta = (float(q["psi"]), float(q["dx"]), float(q["dy"]), float(q["defocus"]))
        s = orc.refine_score(*args, theta)[0]
        g = None
        if i in fd_rows:
            g = np.zeros(4)
            for k, h in enumerate(FD_STEPS):
                tp = np.array(theta)
                tm = tp.copy()
                tp[k] += h
                tm[k] -= h
                g[k] = (orc.refine_score(*args, tp)[0] - orc.refine_score(*args, tm)[0]) / (2 * h)
        return s, g

    with ThreadPoolExecutor(NTHREADS) as ex:
        res = list(ex.map(one, range(len(peaks))))
    return np.array([r[0] for r in res]), {i: r[1] for i, r in enumerate(res) if r[1] is not None}


def _check_refinement(pk, op, df, tms, image, offs, peaks, fd_rows, what):
    white = orc.template_whitened(op, tms)
    tiles = orc.micrograph_prep(op, orc.ctf(df), image)
    got = pk.adjoint(peaks)
    ref, fd = _oracle_objective(op, df, white, tiles, offs, peaks, fd_rows)
    err = np.abs(got["score"] - ref) / np.maximum(np.abs(ref), 1.0)
    print(f"{what}: {len(peaks)} objectives, max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4
    worst = 0.0
    for i, g_fd in fd.items():
        g = np.array([got["d_psi"][i], got["d_dx"][i], got["d_dy"][i], got["d_defocus"][i]],
                     dtype=np.float64)
        tol = 1e-3 * np.maximum(np.abs(g_fd), 0.05 * np.abs(g_fd).max())
        worst = max(worst, float((np.abs(g - g_fd) / tol).max()))
        assert np.all(np.abs(g - g_fd) <= tol), (i, g, g_fd)
    print(f"{what}: {len(fd)} gradients, worst error {worst:.3f} of the tolerance", flush=True)
    start = got["score"]
    refined, fin = pk.refine(peaks, iters=20)
    assert np.all(fin["score"] >= start)
    ref2, _ = _oracle_objective(op, df, white, tiles, offs, refined)
    err2 = np.abs(fin["score"] - ref2) / np.maximum(np.abs(ref2), 1.0)
    print(f"{what}: refined mean objective {start.mean():.3f} -> {fin['score'].mean():.3f}, "
          f"final max rel err {err2.max():.2e}", flush=True)
    assert err2.max() <= 1e-4
    return start, fin["score"]


def test_differentiation_at_the_full_size_shape(gpu):
    import gisspa_amd
    views = _cases.c1_step5_views()
    vsel = views[::103][:16]
    dens = synth.blob_density(60, 180, 1)
    tms = synth.project_templates(dens, 180, vsel)
    df = 1.8
    mg = synth.make_micrograph(5760, 4092, dens, vsel, 300, 0.05, 61, apix=1.0, defocus_um=df,
                               ampconst=0.1, box=180, psi_step=5.0)
    op = orc.params(720, **PARAMS)
    offs = (orc.tile_offsets(5760, 720, 180), orc.tile_offsets(4092, 720, 180))
    with gisspa_amd.Picker(0, **PARAMS) as pk:
        pk.set_templates(tms, df)
        pk.set_micrograph(mg.image, df)
        picks = pk.pick()          # all 72 in-plane steps
        assert len(picks) >= 128
        top = picks[np.argsort(-picks["score"])[:128]]
        base = gisspa_amd.peaks_from_picks(top, PARAMS["phistep"], df)
        peaks = _perturb(base, 5)
        at_grid = pk.adjoint(base)["score"]
        _, final = _check_refinement(pk, op, df, tms, mg.image, offs, peaks, set(range(128)),
                                     "window 720")
    # the perturbation is undone: back to (at least) the grid optimum's level
    assert np.median(final / at_grid) >= 0.999


def test_combined_stress_shape(gpu):
    import gisspa_amd
    box = 256
    views = synth.sphere_eulers(4600)[::575][:8]
    dens = synth.blob_density(60, box, 1)
    tms = synth.project_templates(dens, box, views)
    df = 2.6
    mg = synth.make_micrograph(5760, 4092, dens, views, 300, 0.05, 71, apix=1.0, defocus_um=df,
                               ampconst=0.1, box=box, psi_step=3.0)
    params = dict(PARAMS, phistep=3.0, diameter=0.75 * box, threshold=6.0)
    op = orc.params(720, **params)
    assert orc.num_psi(3.0) == 120
    ref = orc.pick_norm0(op, orc.ctf(df), tms, mg.image, 0, 2, nthreads=NTHREADS)
    with gisspa_amd.Picker(0, **params) as pk:
        pk.set_templates(tms, df)
        pk.set_micrograph(mg.image, df)
        picks = pk.pick(0, 2)
        got = pk.mapstats().reshape(ref.shape)
        _compare_maps(got, ref, "256 px templates, 3 degree steps")
        _compare_picks(picks, orc.select_norm0(op, ref, 0, 5760, 4092), "stress shape")
        top = picks[np.argsort(-picks["score"])[:64]]
        assert len(top) == 64
        peaks = _perturb(gisspa_amd.peaks_from_picks(top, 3.0, df), 6)
        offs = (orc.tile_offsets(5760, 720, 180), orc.tile_offsets(4092, 720, 180))
        _check_refinement(pk, op, df, tms, mg.image, offs, peaks, set(range(0, 64, 4)),
                          "stress shape")
