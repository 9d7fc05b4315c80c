"""GPU parity: the HIP picker against the FP64 oracle on seeded small cases.

Tolerances (FP32 device arithmetic vs FP64 oracle):
  * intermediate arrays: max |gpu - oracle| <= 2e-5 * max |oracle|;
  * per-map scores: |gpu - oracle| <= 1e-4 * max(|oracle|, 1): relative for
    every score that can cross a threshold (>= 1), absolute below, where a
    near-zero peak makes the relative error ill-conditioned;
  * peak positions: identical, except where the oracle's own CC map holds a
    second position within 1e-5 (relative) of the peak (a near tie).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import _cases

pytestmark = pytest.mark.gpu


def _picker(case):
    import gisspa_amd
    pk = gisspa_amd.Picker(0, **case.params)
    pk.set_templates(case.templates, case.defocus)
    pk.set_micrograph(case.image, case.defocus)
    return pk


def _oparams(case):
    return orc.params(case.l, **case.params)


@pytest.fixture(scope="module", params=[0, 1])
def tiny_case(request):
    return _cases.tiny(seed=request.param, l=128 if request.param == 0 else 96)


def test_template_real(gpu, tiny_case):
    c = tiny_case
    with _picker(c) as pk:
        got = pk.template_real()
    ref = orc.template_prep(_oparams(c), orc.ctf(c.defocus), c.templates).real
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("template real rel err", err)
    assert err < 2e-5


def test_tile_spectra(gpu, tiny_case):
    c = tiny_case
    with _picker(c) as pk:
        got = pk.tile_spectra()
    ref = orc.micrograph_prep(_oparams(c), orc.ctf(c.defocus), c.image)
    H = c.l // 2 + 1
    ref_h = np.transpose(ref, (0, 2, 1))[:, :H, :]
    err = np.abs(got - ref_h).max() / np.abs(ref_h).max()
    print("tile spectra rel err", err)
    assert err < 2e-5


def test_rotated_spectra(gpu, tiny_case):
    c = tiny_case
    op = _oparams(c)
    treal = orc.template_prep(op, orc.ctf(c.defocus), c.templates)
    H = c.l // 2 + 1
    with _picker(c) as pk:
        for ip in (0, 1, 5):
            got = pk.rotated_spectra(ip)
            ref = orc.rotate_normalize(op, treal, orc.psi_value(op.phi_step, ip))
            ref_h = np.transpose(ref, (0, 2, 1))[:, :H, :]
            err = np.abs(got - ref_h).max() / np.abs(ref_h).max()
            print("rotated spectra psi", ip, "rel err", err)
            assert err < 2e-5


def test_mapstats_and_picks(gpu, tiny_case):
    c = tiny_case
    op = _oparams(c)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with _picker(c) as pk:
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    rs, gs_ = ref["score"], got["score"].astype(np.float64)
    rel = np.abs(gs_ - rs) / np.maximum(np.abs(rs), 1.0)
    w = np.unravel_index(np.argmax(rel), rel.shape)
    print("score max rel err", rel.max(), "at", w, "gpu", gs_[w], "oracle", rs[w],
          "max score", rs.max())
    assert rel.max() < 1e-4
    mism = np.argwhere(got["pos"] != ref["pos"])
    print("position mismatches", len(mism), "of", ref.size)
    assert len(mism) == 0
    # pick list: identical records in identical order
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    print("picks gpu", len(picks), "oracle", len(sel))
    assert len(picks) == len(sel)
    assert np.array_equal(picks["psi_index"], sel["psi_index"])
    assert np.array_equal(picks["view"], sel["view"])
    assert np.array_equal(picks["cx"], sel["cx"].astype(np.float32))
    assert np.array_equal(picks["cy"], sel["cy"].astype(np.float32))
    assert np.allclose(picks["score"], sel["score"], rtol=1e-4)


N1_CASES = {
    # (case arguments, views per batch: None = the default, one batch per psi)
    "l64": (dict(seed=3, l=64, box=48, n_views=4, size=(200, 176), phistep=45.0, overlap=16,
                 threshold=3.0), None),
    # batches of 2 views over all tiles, the last one ragged (5 views = 2 + 2 + 1)
    "l96-groups": (dict(seed=9, l=96, box=48, n_views=5, size=(256, 224), phistep=60.0,
                        overlap=24, threshold=3.0), 2),
    "l128-one-view": (dict(seed=10, l=128, box=64, n_views=3, size=(320, 288), phistep=90.0,
                           overlap=32, threshold=3.0), 1),
}


@pytest.mark.parametrize("name", list(N1_CASES))
def test_norm_type1_picks(gpu, name):
    """norm_type 1: per-pixel mean/sd over every view x psi, z-scores, and one
    candidate per 1024-pixel run (main.cu:1170-1229).  Batches hold groups of
    views over all tiles (the accumulation sums a tile's views of a batch in
    registers); the cases cover one group per psi, several groups with a
    ragged last one, and one view per group.  Picks must match the oracle's
    record for record; a record whose oracle z-score lies within 1e-4
    (relative) of the threshold may be missing on either side."""
    import gisspa_amd
    args, views_per_batch = N1_CASES[name]
    c = _cases.tiny(**args)
    c.params["norm_type"] = 1
    op = _oparams(c)
    ref = orc.pick_norm1(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        if views_per_batch:
            pk.set_batch(pk.info().n_tiles * views_per_batch)
        got = pk.pick()
        # the accumulations run in batch order on both lane streams: a second
        # search gives the same bits
        again = pk.pick()
    assert np.array_equal(got, again)
    thr = c.params["threshold"]
    key = lambda r: (int(r["psi_index"]), int(r["view"]), float(r["cx"]), float(r["cy"]))
    rk = {key(r): float(r["score"]) for r in ref}
    gk = {key(r): float(r["score"]) for r in got}
    print("norm1 picks gpu", len(got), "oracle", len(ref))
    assert len(ref) > 10
    for k in set(rk) ^ set(gk):
        s = rk.get(k, gk.get(k))
        assert abs(s - thr) <= 1e-4 * thr, (k, s)
    for k in set(rk) & set(gk):
        assert abs(rk[k] - gk[k]) <= 1e-4 * max(1.0, abs(rk[k])), (k, rk[k], gk[k])
    # canonical order: psi, then tile, view, block
    order = [(int(r["psi_index"]), int(r["tile"]), int(r["view"])) for r in got]
    assert order == sorted(order)


@pytest.mark.parametrize("size", [(320, 288), (330, 294)])
def test_phase_flip_path(gpu, size):
    """phase_flip = 1: whole-micrograph CTF sign flip, radial whitening,
    weighting and band pass (init_d_image, main.cu:548-586) on sizes with
    factors 3, 5, 7, 11 (330 = 2 3 5 11, 294 = 2 3 7^2), then per-tile
    amplitude normalisation (main.cu:617-653) and the norm_type 0 picks."""
    import gisspa_amd
    c = _cases.tiny(seed=4, l=96, size=size, threshold=4.0)
    c.params["phase_flip"] = 1
    op = _oparams(c)
    ref_tiles = orc.micrograph_prep(op, orc.ctf(c.defocus), c.image)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        tiles = pk.tile_spectra()
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    H = c.l // 2 + 1
    rh = np.transpose(ref_tiles, (0, 2, 1))[:, :H, :]
    err = np.abs(tiles - rh).max() / np.abs(rh).max()
    print("phase-flip tile spectra rel err", err)
    assert err < 2e-5
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("phase-flip score max rel err", rel.max())
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"])


@pytest.mark.parametrize("phase_flip", [0, 1])
def test_invert_path(gpu, phase_flip):
    """invert = 1: the micrograph's contrast inverted after the hot-pixel
    clamp (init_d_image, main.cu:523-547), with and without the phase flip;
    tile spectra, per-map scores and picks against the oracle."""
    import gisspa_amd
    c = _cases.tiny(seed=11, l=96, threshold=3.5)
    c.params["invert"] = 1
    c.params["phase_flip"] = phase_flip
    op = _oparams(c)
    ref_tiles = orc.micrograph_prep(op, orc.ctf(c.defocus), c.image)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        tiles = pk.tile_spectra()
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    H = c.l // 2 + 1
    rh = np.transpose(ref_tiles, (0, 2, 1))[:, :H, :]
    err = np.abs(tiles - rh).max() / np.abs(rh).max()
    print("invert tile spectra rel err", err)
    assert err < 2e-5
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("invert score max rel err", rel.max(), "picks", len(picks))
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"])
    assert np.array_equal(sel["cx"].astype(np.float32), picks["cx"])


def test_window_adjusted_to_the_templates(gpu):
    """Templates larger than the configured window: the window grows to the
    smallest multiple of 32 that holds them (adjustPaddingSize,
    main.cu:158-178; 80 px templates, window 64 -> 96), while the CTF
    sampling of the template weighting keeps the configured window
    (ds, main.cu:155 before main.cu:1116).  The overlap key is left out, so
    tiles use the reference's default stride l + 1 (main.cu:228)."""
    import gisspa_amd
    c = _cases.tiny(seed=12, l=64, box=80, n_views=4, size=(352, 320), threshold=3.5)
    del c.params["overlap"]
    op = orc.params(96, **c.params)
    assert op.l == 96 and op.l_config == 64 and op.overlap == -1
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        info = pk.info()
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    assert info.l == 96
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("adjusted window: score max rel err", rel.max(), "maps", ref.size, "picks", len(picks))
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"])
    assert np.array_equal(sel["cx"].astype(np.float32), picks["cx"])


@pytest.mark.parametrize("phase_flip", [0, 1])
def test_astigmatic_ctf(gpu, phase_flip):
    """Astigmatic CTF (dfdiff 0.12 um, dfang 37 degrees; CTF_AST,
    GPU_func.cu:292-315) in the template weighting and, with phase_flip 1,
    in the micrograph's flip and weighting, on a non-square micrograph:
    template real parts, tile spectra, per-map scores and picks against the
    oracle."""
    import gisspa_amd
    c = _cases.tiny(seed=13, l=96, size=(352, 256), threshold=3.5)
    c.params["phase_flip"] = phase_flip
    op = _oparams(c)
    ctf = orc.ctf(c.defocus, 0.12, 37.0)
    ref_t = orc.template_prep(op, ctf, c.templates)
    ref_tiles = orc.micrograph_prep(op, ctf, c.image)
    ref = orc.pick_norm0(op, ctf, c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus, 0.12, 37.0)
        treal = pk.template_real()
        pk.set_micrograph(c.image, c.defocus, 0.12, 37.0)
        tiles = pk.tile_spectra()
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    et = np.abs(treal - ref_t.real).max() / np.abs(ref_t.real).max()
    H = c.l // 2 + 1
    rh = np.transpose(ref_tiles, (0, 2, 1))[:, :H, :]
    es = np.abs(tiles - rh).max() / np.abs(rh).max()
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("astigmatic: template", et, "tiles", es, "scores", rel.max(), "picks", len(picks))
    assert et < 2e-5 and es < 2e-5
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"])


@pytest.mark.parametrize("l,size,box,overlap", [
    (96, (80, 72), 32, 24),          # micrograph smaller than the window: one zero-padded tile
    (1024, (1500, 1100), 128, 256),  # the largest compiled window
])
def test_window_edge_cases(gpu, l, size, box, overlap):
    import gisspa_amd
    c = _cases.tiny(seed=14, l=l, box=box, n_views=2, size=size, phistep=180.0,
                    overlap=overlap, threshold=3.0)
    op = _oparams(c)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print(f"window {l}, micrograph {size}: {ref.shape} maps, score max rel err {rel.max():.2e}, "
          f"picks {len(picks)}")
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) and np.array_equal(sel["view"], picks["view"])


def test_norm_type1_batch_of_micrographs(gpu):
    """norm_type 1 over two micrographs of different sizes in one tile pool
    (gs_set_micrographs): the per-pixel statistics are per tile, so each
    micrograph's records equal the oracle's for that micrograph alone, and
    records come micrograph by micrograph."""
    import gisspa_amd
    c1 = _cases.tiny(seed=3, l=64, box=48, n_views=4, size=(200, 176), phistep=45.0,
                     overlap=16, threshold=3.0)
    c2 = _cases.tiny(seed=15, l=64, box=48, n_views=4, size=(176, 232), phistep=45.0,
                     overlap=16, threshold=3.0)
    c1.params["norm_type"] = 1
    op = _oparams(c1)
    thr = c1.params["threshold"]
    with gisspa_amd.Picker(0, **c1.params) as pk:
        pk.set_templates(c1.templates, c1.defocus)
        pk.set_micrographs([c1.image, c2.image], [c1.defocus, c1.defocus])
        got = pk.pick()
    assert np.all(np.diff(got["mic"]) >= 0)
    key = lambda r: (int(r["psi_index"]), int(r["view"]), float(r["cx"]), float(r["cy"]))
    for i, c in enumerate((c1, c2)):
        ref = orc.pick_norm1(op, orc.ctf(c1.defocus), c1.templates, c.image)
        mine = got[got["mic"] == i]
        rk = {key(r): float(r["score"]) for r in ref}
        gk = {key(r): float(r["score"]) for r in mine}
        print("mic", i, "norm1 records gpu", len(gk), "oracle", len(rk))
        assert len(rk) > 10
        for k in set(rk) ^ set(gk):
            s = rk.get(k, gk.get(k))
            assert abs(s - thr) <= 1e-4 * thr, (i, k, s)
        for k in set(rk) & set(gk):
            assert abs(rk[k] - gk[k]) <= 1e-4 * max(1.0, abs(rk[k])), (i, k)
