"""gs_remove_repeats_noalloc (the allocation-free entry point, static device
scratch) against gs_remove_repeats (its own allocations), both through the
C ABI: identical keep flags and counts on two calls in a row, the second
with a smaller list, and a third list whose undecided pairs go to the
host's acos; a list above the pick limit is refused before any device work."""
import ctypes

import numpy as np
import pytest

from tests.test_gpu_dedup import _picks

pytestmark = pytest.mark.gpu

GS_ERR_CAPACITY = -6


def _call(fn, cols, center, thr):
    import gisspa_amd
    recs = gisspa_amd.pick_records(cols["group"], cols["x"], cols["y"], cols["alt"],
                                   cols["az"], cols["phi"], cols["score"])
    recs = np.ascontiguousarray(recs)
    keep = np.full(len(recs), 7, dtype=np.uint8)
    nk = ctypes.c_size_t(12345)
    rc = fn(0, recs.ctypes.data, len(recs), float(center), float(thr), keep.ctypes.data,
            ctypes.byref(nk))
    return rc, keep, nk.value


def _ties(n=40):
    return {"group": np.zeros(n, int), "x": np.full(n, 100.0), "y": np.full(n, 200.0),
            "alt": np.full(n, 10.0), "az": np.full(n, 20.0),
            "phi": 5.0 * np.arange(n) % 360, "score": np.linspace(9, 7, n)}


def test_noalloc_matches_allocating_entry_point(gpu):
    import gisspa_amd
    lib = gisspa_amd.load()
    cases = [(_picks(3, [9000, 6000, 5000], spread=8.0), 12.0, 10.0),   # 20,000 picks
             (_picks(0, [300, 1, 57, 800]), 10.0, 8.0),                 # 1,158 picks
             (_ties(), 10.0, 5.0)]                                      # acos on the host
    for k, (cols, center, thr) in enumerate(cases):
        print(f"case {k}: {len(cols['x'])} picks, center {center}, thr {thr}", flush=True)
        rc0, keep0, n0 = _call(lib.gs_remove_repeats, cols, center, thr)
        rc1, keep1, n1 = _call(lib.gs_remove_repeats_noalloc, cols, center, thr)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        assert np.array_equal(keep0, keep1), np.count_nonzero(keep0 != keep1)
        assert n0 == n1 == int(keep0.sum())
        assert 0 < n0 < len(keep0)


def test_noalloc_rejects_too_many_picks(gpu):
    import gisspa_amd
    lib = gisspa_amd.load()
    n = 65536 + 1
    cols = {"group": np.zeros(n, int), "x": np.arange(n, dtype=float), "y": np.zeros(n),
            "alt": np.zeros(n), "az": np.zeros(n), "phi": np.zeros(n), "score": np.ones(n)}
    rc, keep, nk = _call(lib.gs_remove_repeats_noalloc, cols, 5.0, 5.0)
    assert rc == GS_ERR_CAPACITY
    assert nk == 0
