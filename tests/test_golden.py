"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py): a
seeded small case's inputs, the FP64 oracle's per-map results and pick list,
and the output lines of that pick list (main.cu:985 format).

CPU: the generator and the oracle reproduce the fixture exactly (drift of
either fails here).  GPU: the picker and the ./main CLI on the fixture's
inputs give the fixture's results (scores within 1e-4 * max(|s|, 1),
identical positions, identical pick list, the same lines with scores to
1e-4)."""
import os
import re
import subprocess

import numpy as np
import pytest

from gisspa_amd import synth
from oracle import oracle as orc
from tests import _cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LINE = re.compile(r"^(\d+)\t(\S+)\tdefocus=([-\d.]+)\tdfdiff=([-\d.]+)\tdfang=([-\d.]+)\t"
                  r"euler=([-\d.]+),([-\d.]+),([-\d.]+)\tcenter=([-\d.]+),([-\d.]+)\tscore=([-\d.]+)$")


@pytest.fixture(scope="module")
def golden():
    d = np.load(os.path.join(HERE, "tiny_seed0.npz"))
    with open(os.path.join(HERE, "tiny_seed0_picks.lst")) as f:
        lines = f.read().splitlines()
    return d, lines


def test_generator_reproduces_the_inputs(golden):
    d, _ = golden
    c = _cases.tiny(seed=0, l=128)
    assert np.array_equal(c.templates, d["templates"])
    assert np.array_equal(c.image, d["image"])


def test_oracle_reproduces_the_results(golden):
    d, lines = golden
    c = _cases.tiny(seed=0, l=128)
    p = orc.params(c.l, **c.params)
    st = orc.pick_norm0(p, orc.ctf(float(d["defocus"])), d["templates"], d["image"])
    ref = d["mapstats"]
    assert np.array_equal(st["pos"], ref["pos"])
    assert np.allclose(st["score"], ref["score"], rtol=1e-12, atol=1e-12)
    sel = orc.select_norm0(p, st, 0, d["image"].shape[1], d["image"].shape[0])
    assert np.array_equal(sel, d["picks"])
    assert len(lines) == len(sel)


@pytest.mark.gpu
def test_picker_matches_the_fixture(gpu, golden):
    import gisspa_amd
    d, _ = golden
    c = _cases.tiny(seed=0, l=128)
    ref = d["mapstats"]
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates(d["templates"], float(d["defocus"]))
        pk.set_micrograph(d["image"], float(d["defocus"]))
        picks = pk.pick()
        got = pk.mapstats().reshape(ref.shape)
    rel = np.abs(got["score"] - ref["score"]) / np.maximum(np.abs(ref["score"]), 1.0)
    print("golden: score max rel err", rel.max(), "picks", len(picks))
    assert rel.max() < 1e-4
    assert np.array_equal(got["pos"], ref["pos"])
    sel = d["picks"]
    assert len(picks) == len(sel)
    assert np.array_equal(picks["view"], sel["view"])
    assert np.array_equal(picks["psi_index"], sel["psi_index"])
    assert np.array_equal(picks["cx"], sel["cx"].astype(np.float32))


@pytest.mark.gpu
def test_cli_prints_the_fixture_lines(gpu, golden, tmp_path):
    from gisspa_amd import MAIN_PATH
    d, lines = golden
    c = _cases.tiny(seed=0, l=128)
    eul = synth.random_eulers(d["templates"].shape[0], 12)
    wl = synth.Workload("golden", d["templates"], eul, [(d["image"], float(d["defocus"]))],
                        dict(c.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = open(os.path.join(tmp_path, "picks.lst")).read().splitlines()
    assert len(out) == len(lines)
    for a, b in zip(out, lines):
        ma, mb = LINE.match(a), LINE.match(b)
        assert ma and mb, (a, b)
        assert ma.group(2).endswith(mb.group(2))
        for k in range(3, 11):   # defocus, dfdiff, dfang, euler, centre: as printed
            assert ma.group(k) == mb.group(k), (k, a, b)
        assert float(ma.group(11)) == pytest.approx(float(mb.group(11)), rel=1e-4, abs=2e-6)
