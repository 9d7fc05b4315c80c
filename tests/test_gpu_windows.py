"""GPU parity at every family of window size the reference can produce.

adjustPaddingSize (main.cu:158-178) grows a window smaller than the
templates to the next multiple of 32, so any multiple of 32 can occur.  The
transforms of length L and L/2 then need radices with prime factors up to 31
(224 = 7 * 32, 352 = 11 * 32, ..., 992 = 31 * 32).  One size per prime
factor (and the odd powers 9, 25, 27) runs the whole picker against the
oracle: template real parts and tile spectra within 2e-5, per-map scores
within 1e-4 * max(|s|, 1), identical peak positions and pick lists.  Also at
a prime-factor window: norm_type 1 picks, and the refinement objective and
gradient against FP64 central differences.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import _cases

pytestmark = pytest.mark.gpu


def _run(c, views=None):
    import gisspa_amd
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        treal = pk.template_real()
        pk.set_micrograph(c.image, c.defocus)
        tiles = pk.tile_spectra()
        picks = pk.pick()
        st = pk.mapstats()
        info = pk.info()
    return treal, tiles, picks, st, info


# window -> prime factor of L / 32 that the transforms exercise
WINDOWS = {224: 7, 288: 9, 352: 11, 416: 13, 544: 17, 608: 19, 736: 23, 800: 25, 864: 27,
           928: 29, 992: 31}


@pytest.mark.parametrize("l", list(WINDOWS))
def test_window_size_matches_oracle(gpu, l):
    box = 64 if l < 600 else 96
    size = (l + l // 2 + 10, l + l // 3 + 6)
    c = _cases.tiny(seed=30 + l % 17, l=l, box=box, n_views=2, size=size, phistep=90.0,
                    overlap=l // 4, threshold=3.0)
    op = orc.params(l, **c.params)
    print(f"window {l} (factor {WINDOWS[l]}), micrograph {size}", flush=True)
    ref_t = orc.template_prep(op, orc.ctf(c.defocus), c.templates)
    ref_tiles = orc.micrograph_prep(op, orc.ctf(c.defocus), c.image)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    treal, tiles, picks, st, info = _run(c)
    assert info.l == l
    et = np.abs(treal - ref_t.real).max() / np.abs(ref_t.real).max()
    H = l // 2 + 1
    rh = np.transpose(ref_tiles, (0, 2, 1))[:, :H, :]
    es = np.abs(tiles - rh).max() / np.abs(rh).max()
    got = st.reshape(ref.shape)
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print(f"  {ref.size} maps: template {et:.1e}, tiles {es:.1e}, scores {rel.max():.1e}, "
          f"picks {len(picks)}")
    assert et < 2e-5 and es < 2e-5
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks) > 0
    assert np.array_equal(sel["view"], picks["view"])
    assert np.array_equal(sel["cx"].astype(np.float32), picks["cx"])


def test_window_grown_from_180_to_224(gpu):
    """200-px templates with window_size 180: the window grows to 224
    (main.cu:158-178) and the weighting keeps the configured window's
    frequency step (main.cu:155)."""
    c = _cases.tiny(seed=41, l=180, box=200, n_views=2, size=(480, 420), phistep=120.0,
                    overlap=60, threshold=3.0)
    op = orc.params(224, **c.params)
    assert op.l == 224 and op.l_config == 180
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    treal, tiles, picks, st, info = _run(c)
    assert info.l == 224
    got = st.reshape(ref.shape)
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print(f"grown window: {ref.size} maps, scores {rel.max():.1e}, picks {len(picks)}")
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(sel) == len(picks)
    assert np.array_equal(sel["view"], picks["view"])


def test_norm_type1_at_window_352(gpu):
    import gisspa_amd
    c = _cases.tiny(seed=42, l=352, box=64, n_views=3, size=(560, 480), phistep=120.0,
                    overlap=88, threshold=2.5)
    c.params["norm_type"] = 1
    op = orc.params(352, **c.params)
    ref = orc.pick_norm1(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        got = pk.pick()
    thr = c.params["threshold"]
    key = lambda r: (int(r["psi_index"]), int(r["view"]), float(r["cx"]), float(r["cy"]))
    rk = {key(r): float(r["score"]) for r in ref}
    gk = {key(r): float(r["score"]) for r in got}
    print("norm1 at 352: picks gpu", len(got), "oracle", len(ref))
    assert len(ref) > 10
    for k in set(rk) ^ set(gk):
        s = rk.get(k, gk.get(k))
        assert abs(s - thr) <= 1e-4 * thr, (k, s)
    for k in set(rk) & set(gk):
        assert abs(rk[k] - gk[k]) <= 1e-4 * max(1.0, abs(rk[k])), (k, rk[k], gk[k])


def test_refinement_gradient_at_window_416(gpu):
    """The adjoint at a window whose transforms use radix 13: objective within
    1e-4 and gradient within 1e-3 of FP64 central differences."""
    import gisspa_amd
    from tests.test_gpu_refine import FD_STEPS
    c = _cases.tiny(seed=43, l=416, box=64, n_views=3, size=(520, 480), phistep=30.0,
                    overlap=104, threshold=3.0)
    op = orc.params(416, **c.params)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
        top = picks[np.argsort(-picks["score"])[:6]]
        base = gisspa_amd.peaks_from_picks(top, c.params["phistep"], c.defocus)
        rng = np.random.default_rng(7)
        p = base.copy()
        p["psi"] += rng.uniform(-2.5, 2.5, len(p)).astype(np.float32)
        p["dx"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
        p["dy"] += rng.uniform(-1, 1, len(p)).astype(np.float32)
        p["defocus"] += rng.uniform(-0.02, 0.02, len(p)).astype(np.float32)
        got = pk.adjoint(p)
    white = orc.template_whitened(op, c.templates)
    tiles = orc.micrograph_prep(op, orc.ctf(c.defocus), c.image)
    ny, nx = c.image.shape
    ox = orc.tile_offsets(nx, 416, c.params["overlap"])
    oy = orc.tile_offsets(ny, 416, c.params["overlap"])
    for i, q in enumerate(p):
        t = int(q["tile"])
        args = (op, orc.ctf(c.defocus), white[int(q["view"])], tiles[t],
                float(q["cx"]) - ox[t % len(ox)], float(q["cy"]) - oy[t // len(ox)])
        theta = (float(q["psi"]), float(q["dx"]), float(q["dy"]), float(q["defocus"]))
        ref = orc.refine_score(*args, theta)[0]
        assert abs(got["score"][i] - ref) <= 1e-4 * max(abs(ref), 1.0)
        fd = np.zeros(4)
        for k, h in enumerate(FD_STEPS):
            tp = np.array(theta)
            tm = tp.copy()
            tp[k] += h
            tm[k] -= h
            fd[k] = (orc.refine_score(*args, tp)[0] - orc.refine_score(*args, tm)[0]) / (2 * h)
        g = np.array([got["d_psi"][i], got["d_dx"][i], got["d_dy"][i], got["d_defocus"][i]])
        tol = 1e-3 * np.maximum(np.abs(fd), 0.05 * np.abs(fd).max())
        print(i, "grad", g, "fd", fd)
        assert np.all(np.abs(g - fd) <= tol), (i, g, fd)
