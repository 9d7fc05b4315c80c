"""Duplicate removal on the GPU (gs_remove_repeats, the remove_repeats tool)
against the FP64 oracle of python_scripts/remove_repeat_particles_from_list.py:
the keep flags must be identical.  Inputs cover clustered picks in groups of
uneven size (one of a single pick), tied scores, identical orientations,
180-degree turns, tilt flips, centre distances exactly at the threshold and
in-plane steps exactly at the Euler threshold (where acos rounding decides and
the library settles the pair with the C library's acos)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _picks(seed, sizes, step=5.0, spread=4.0):
    rng = np.random.default_rng(seed)
    cols = {k: [] for k in ("group", "x", "y", "alt", "az", "phi", "score")}
    views = np.array([[0, 0], [180, 0], [10, 20], [10, 25], [90, 45], [45.5, 300]])
    for g, n in enumerate(sizes):
        nc = max(1, n // 12)
        centres = rng.uniform(0, 4000, (nc, 2))
        c = rng.integers(0, nc, n)
        xy = centres[c] + np.round(rng.normal(0, spread, (n, 2)), 2)
        v = rng.integers(0, len(views), n)
        cols["group"] += [g] * n
        cols["x"] += list(xy[:, 0])
        cols["y"] += list(xy[:, 1])
        cols["alt"] += list(views[v, 0])
        cols["az"] += list(views[v, 1])
        cols["phi"] += list(step * rng.integers(0, 72, n))
        cols["score"] += list(np.round(rng.uniform(7, 9, n), 1))   # many ties
    return {k: np.array(v) for k, v in cols.items()}


def _compare(cols, center, thr):
    import gisspa_amd
    recs = gisspa_amd.pick_records(cols["group"], cols["x"], cols["y"], cols["alt"],
                                   cols["az"], cols["phi"], cols["score"])
    got = gisspa_amd.remove_repeats(recs, center, thr)
    ref = orc.remove_repeats(cols["group"], cols["x"], cols["y"], cols["alt"], cols["az"],
                             cols["phi"], cols["score"], center, thr)
    print(f"n={len(ref)} kept={ref.sum()} center={center} thr={thr} "
          f"mismatches={np.count_nonzero(got != ref)}")
    assert np.array_equal(got, ref)
    return ref


@pytest.mark.parametrize("center,thr", [(10.0, 8.0), (15.0, 5.0), (6.0, 10.0), (0.0, 5.0)])
def test_matches_oracle(gpu, center, thr):
    cols = _picks(0, [300, 1, 57, 800])
    ref = _compare(cols, center, thr)
    if center > 0:
        assert 0 < ref.sum() < len(ref)


def test_threshold_ties(gpu):
    # same place and view, in-plane angles exactly one threshold apart: the
    # distance 5.00000000000016 (not < 5) comes from acos's last bits
    n = 40
    cols = {"group": np.zeros(n, int), "x": np.full(n, 100.0), "y": np.full(n, 200.0),
            "alt": np.full(n, 10.0), "az": np.full(n, 20.0),
            "phi": 5.0 * np.arange(n) % 360, "score": np.linspace(9, 7, n)}
    _compare(cols, 10.0, 5.0)
    # the centre distance exactly at the threshold (3-4-5)
    cols["x"] = 100.0 + 3.0 * (np.arange(n) % 2)
    cols["y"] = 200.0 + 4.0 * (np.arange(n) % 2)
    _compare(cols, 5.0, 12.0)


def test_large_list(gpu):
    cols = _picks(3, [9000, 6000, 5000], spread=8.0)
    _compare(cols, 12.0, 10.0)


def test_cli_matches_oracle(gpu, tmp_path):
    import gisspa_amd
    tool = os.path.join(os.path.dirname(gisspa_amd.LIB_PATH), "remove_repeats")
    cols = _picks(5, [200, 150], step=5.0)
    paths = ["mics/a.mrc", "mics/b.mrc"]
    lines = []
    for i in range(len(cols["x"])):
        g = int(cols["group"][i])
        lines.append(f"{i}\t{paths[g]}\tdefocus=1.800000\tdfdiff=0.000000\tdfang=0.000000\t"
                     f"euler={cols['alt'][i]:f},{cols['az'][i]:f},{cols['phi'][i]:f}\t"
                     f"center={cols['x'][i]:f},{cols['y'][i]:f}\tscore={cols['score'][i]:f}\n")
    src = tmp_path / "picks.lst"
    src.write_text("".join(lines))
    out = tmp_path / "unique.lst"
    r = subprocess.run([tool, str(src), "2", "10", "8", str(out)], capture_output=True,
                       text=True, timeout=120)
    print(r.stderr)
    assert r.returncode == 0
    # the oracle on the values as printed
    pr = lambda v: np.array([float(f"{x:f}") for x in v])
    ref = orc.remove_repeats(cols["group"], pr(cols["x"]), pr(cols["y"]), pr(cols["alt"]),
                             pr(cols["az"]), pr(cols["phi"]), pr(cols["score"]), 10.0, 8.0)
    expect = "".join(l for l, k in zip(lines, ref) if k)
    assert out.read_text() == expect
    assert 0 < ref.sum() < len(ref)
