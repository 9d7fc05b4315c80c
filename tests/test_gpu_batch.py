"""GPU parity of the batched, plain and graph-replay paths.

  * a batch of micrographs of different sizes and CTFs (gs_set_micrographs)
    gives, micrograph by micrograph, the oracle's per-map scores and the
    oracle's pick list (the reference picks one micrograph per loop
    iteration, main.cu:1149-1269);
  * the stage-by-stage mode (GS_MODE_PLAIN) gives the oracle's scores;
  * per-psi HIP graphs replay to the same bits as direct launches.

Tolerances as in test_gpu_parity: scores within 1e-4 * max(|oracle|, 1),
identical peak positions.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import _cases

pytestmark = pytest.mark.gpu


def _check_scores(got, ref, what):
    rel = np.abs(got["score"].astype(np.float64) - ref["score"]) / np.maximum(
        np.abs(ref["score"]), 1.0)
    print(what, "score max rel err", rel.max())
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"]), what


def _oracle_maps(op, tctf, mctf, tm, img):
    """[psi][tile][view] scores with the templates weighted by the first
    micrograph's CTF and each micrograph phase-flipped by its own (the
    reference prepares the templates once, from the first micrograph's
    header, main.cu:1104-1129, and reads each micrograph's CTF in the loop,
    main.cu:1159-1163)."""
    treal = orc.template_prep(op, tctf, tm)
    tiles = orc.micrograph_prep(op, mctf, img)
    P = orc.num_psi(op.phi_step)
    return np.stack([orc.correlate_many(op, tiles, orc.rotate_normalize(
        op, treal, orc.psi_value(op.phi_step, ip))) for ip in range(P)])


def _batch_cases():
    sizes = [(320, 288), (256, 352), (400, 200)]
    out = []
    for i, sz in enumerate(sizes):
        c = _cases.tiny(seed=20 + i, l=96, size=sz, threshold=4.0)
        out.append(c)
    # one template set (case 0's), three micrographs with their own defocus
    return out


@pytest.mark.parametrize("phase_flip", [0, 1])
def test_batch_of_micrographs(gpu, phase_flip):
    import gisspa_amd
    cases = _batch_cases()
    c0 = cases[0]
    params = dict(c0.params, phase_flip=phase_flip)
    defs = [1.5, 1.2, 1.9]
    op = orc.params(c0.l, **params)
    with gisspa_amd.Picker(0, **params) as pk:
        pk.set_templates(c0.templates, defs[0])
        pk.set_micrographs([c.image for c in cases], defs)
        info = pk.info()
        picks = pk.pick()
        st = pk.mapstats()
    assert info.n_mics == 3
    P, N = info.n_psi, info.n_views
    st = st.reshape(P, info.n_tiles, N)
    t0 = 0
    for i, c in enumerate(cases):
        ny, nx = c.image.shape
        ref = _oracle_maps(op, orc.ctf(defs[0]), orc.ctf(defs[i]), c0.templates, c.image)
        nt = ref.shape[1]
        got = st[:, t0:t0 + nt, :]
        t0 += nt
        _check_scores(got, ref, f"mic {i} ({nx}x{ny})")
        sel = orc.select_norm0(op, ref, 0, nx, ny)
        mine = picks[picks["mic"] == i]
        print("mic", i, "picks gpu", len(mine), "oracle", len(sel))
        assert len(mine) == len(sel)
        assert np.array_equal(mine["psi_index"], sel["psi_index"])
        assert np.array_equal(mine["view"], sel["view"])
        assert np.array_equal(mine["cx"], sel["cx"].astype(np.float32))
        assert np.array_equal(mine["cy"], sel["cy"].astype(np.float32))
    assert t0 == info.n_tiles
    # one micrograph after the other
    assert np.all(np.diff(picks["mic"]) >= 0)


def test_plain_mode_matches_oracle(gpu):
    import gisspa_amd
    c = _cases.tiny(seed=5, l=128)
    op = orc.params(c.l, **c.params)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_mode(gisspa_amd.GS_MODE_PLAIN)
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    _check_scores(got, ref, "plain")
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(picks) == len(sel) and np.array_equal(picks["view"], sel["view"])


def test_graph_replay_is_bit_exact(gpu):
    import gisspa_amd
    c = _cases.tiny(seed=6, l=96)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(c.templates, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        base_picks = pk.pick()
        base = pk.mapstats()
        pk.set_graphs(True)
        for rep in range(3):   # capture, then two replays
            picks = pk.pick()
            st = pk.mapstats()
            assert np.array_equal(st, base), f"replay {rep}"
            assert np.array_equal(picks, base_picks), f"replay {rep}"
        # a new micrograph into the same buffers: replays must see it
        c2 = _cases.tiny(seed=7, l=96)
        pk.set_micrograph(c2.image, c2.defocus)
        g = pk.pick()
        gst = pk.mapstats()
        pk.set_graphs(False)
        d = pk.pick()
        dst = pk.mapstats()
    assert np.array_equal(gst, dst)
    assert np.array_equal(g, d)


@pytest.mark.parametrize("lag,ring", [(4, 8), (2, 3), (1, 2)])
def test_pipelines_agree(gpu, lag, ring):
    """The one-launch-per-psi schedule (in-kernel dependencies through a ring
    of intermediates; small rings force slot reuse) and the two-stream
    schedule and the one-stream schedule give the same bits as the
    launch-per-batch schedule."""
    import gisspa_amd
    c = _cases.tiny(seed=8, l=96)
    out = {}
    for mode in (gisspa_amd.GS_PIPE_LAUNCH, gisspa_amd.GS_PIPE_STREAM, gisspa_amd.GS_PIPE_LANES,
                 gisspa_amd.GS_PIPE_SERIAL):
        with gisspa_amd.Picker(0, **c.params) as pk:
            pk.set_pipeline(mode, lag, ring)
            pk.set_batch(5)
            pk.set_templates(c.templates, c.defocus)
            pk.set_micrograph(c.image, c.defocus)
            out[mode] = (pk.pick(), pk.mapstats())
    a = out[gisspa_amd.GS_PIPE_LAUNCH]
    for b in (out[gisspa_amd.GS_PIPE_STREAM], out[gisspa_amd.GS_PIPE_LANES],
              out[gisspa_amd.GS_PIPE_SERIAL]):
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[0], b[0])


def test_mixed_template_boxes(gpu):
    """Templates of boxes 64, 48 and 57 in one set: each is centred in the
    window by its own size (main.cu:206-207), then picked like any set; the
    oracle gets the same templates centred in the window on the host."""
    import gisspa_amd
    c = _cases.tiny(seed=6, l=128, box=64, n_views=6)
    boxes = [64, 48, 57, 64, 48, 57]
    tms = []
    for t, b in zip(c.templates, boxes):
        o = (64 - b) // 2
        tms.append(np.ascontiguousarray(t[o:o + b, o:o + b]))
    L = c.l
    padded = np.zeros((len(tms), L, L), np.float32)
    for j, t in enumerate(tms):
        b = t.shape[0]
        o = (L - b) // 2
        padded[j, o:o + b, o:o + b] = t
    op = orc.params(L, **c.params)
    ref = orc.pick_norm0(op, orc.ctf(c.defocus), padded, c.image)
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates_mixed(tms, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("mixed boxes: score max rel err", rel.max(), "picks", len(picks))
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = orc.select_norm0(op, ref, 0, c.image.shape[1], c.image.shape[0])
    assert len(picks) == len(sel)
    assert np.array_equal(picks["view"], sel["view"])


@pytest.mark.parametrize("tile_block", [1, 5, 1000])
def test_tile_block_order_is_invisible(gpu, tile_block):
    """The map order (tile blocks of any size, the default cache-sized one)
    changes which maps share a launch, never a map's result: statistics and
    picks are bit-identical."""
    import gisspa_amd
    cases = _batch_cases()
    c0 = cases[0]
    defs = [1.5, 1.2, 1.9]

    def run(tb):
        with gisspa_amd.Picker(0, **c0.params) as pk:
            if tb is not None:
                pk.set_tile_block(tb)
            pk.set_batch(37)   # batches that cut through tile blocks
            pk.set_templates(c0.templates, defs[0])
            pk.set_micrographs([c.image for c in cases], defs)
            return pk.pick(), pk.mapstats()

    p0, s0 = run(None)
    p1, s1 = run(tile_block)
    assert np.array_equal(s0, s1)
    assert np.array_equal(p0, p1)


def test_launch_schedule_with_profiling(gpu):
    """The launch-per-batch schedule with per-launch profiling on, in batches
    of 189 maps (480 maps: two full batches and a ragged one), gives the
    two-stream schedule's bits; the profiling events of every schedule are
    created on first use, however many launches a psi takes."""
    import gisspa_amd
    c = _cases.tiny(seed=8, l=96, n_views=24)
    out = {}
    for mode in (gisspa_amd.GS_PIPE_LANES, gisspa_amd.GS_PIPE_LAUNCH, gisspa_amd.GS_PIPE_STREAM,
                 gisspa_amd.GS_PIPE_SERIAL):
        with gisspa_amd.Picker(0, **c.params) as pk:
            pk.set_pipeline(mode)
            pk.set_batch(189)
            pk.set_profiling(True)
            pk.set_templates(c.templates, c.defocus)
            pk.set_micrograph(c.image, c.defocus)
            info = pk.info()
            assert info.n_tiles * info.n_views == 480
            out[mode] = (pk.pick(), pk.mapstats(), pk.timing())
    a = out[gisspa_amd.GS_PIPE_LANES]
    for mode in (gisspa_amd.GS_PIPE_LAUNCH, gisspa_amd.GS_PIPE_STREAM, gisspa_amd.GS_PIPE_SERIAL):
        b = out[mode]
        assert np.array_equal(a[1], b[1]), mode
        assert np.array_equal(a[0], b[0]), mode
        assert b[2].corr_launches >= 1
    assert a[2].corr_launches == 3 and out[gisspa_amd.GS_PIPE_LAUNCH][2].corr_launches == 4
    assert out[gisspa_amd.GS_PIPE_SERIAL][2].corr_launches == 3


@pytest.mark.parametrize("norm_type", [0, 1])
def test_record_buffer_grows(gpu, norm_type):
    """A device-side record buffer smaller than the selection (capacity 1 or
    7 records) is grown to the exact count and the selection runs again: the
    picks equal those of the default capacity, also on a second search in the
    same context."""
    import gisspa_amd
    c = _cases.tiny(seed=5, l=96, threshold=3.0 if norm_type else 4.0)
    c.params["norm_type"] = norm_type
    runs = {}
    for cap in (0, 1, 7):
        with gisspa_amd.Picker(0, **c.params) as pk:
            if cap:
                pk.set_record_capacity(cap)
            pk.set_templates(c.templates, c.defocus)
            pk.set_micrograph(c.image, c.defocus)
            runs[cap] = pk.pick()
            if cap == 7:
                assert np.array_equal(pk.pick(), runs[cap])
    print("norm_type", norm_type, "picks", len(runs[0]))
    assert len(runs[0]) > 7
    for cap in (1, 7):
        assert np.array_equal(runs[cap], runs[0]), cap
