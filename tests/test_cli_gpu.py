"""The ./main CLI end to end on the GPU against the oracle's pick list.

A seeded small workload is written in the reference's formats (config,
micrograph .lst, Euler .lst, MRC stacks); the output file must hold the
oracle's picks, line for line, in the reference's format (main.cu:985),
with scores equal to 1e-4 relative."""
import os
import re
import subprocess

import numpy as np
import pytest

from gisspa_amd import MAIN_PATH, synth
from oracle import oracle as orc
from tests import _cases

pytestmark = pytest.mark.gpu

LINE = re.compile(r"^(\d+)\t(\S+)\tdefocus=([-\d.]+)\tdfdiff=([-\d.]+)\tdfang=([-\d.]+)\t"
                  r"euler=([-\d.]+),([-\d.]+),([-\d.]+)\tcenter=([-\d.]+),([-\d.]+)\tscore=([-\d.]+)$")


def test_cli_matches_oracle(gpu, tmp_path):
    c = _cases.tiny(seed=2, l=128, n_views=5, phistep=45.0, threshold=4.5)
    eul = synth.random_eulers(len(c.templates), 7)
    wl = synth.Workload("cli", c.templates, eul, [(c.image, c.defocus)], dict(c.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = open(os.path.join(tmp_path, "picks.lst")).read().splitlines()
    op = orc.params(c.l, **c.params)
    st = orc.pick_norm0(op, orc.ctf(c.defocus), c.templates, c.image)
    sel = orc.select_norm0(op, st, 0, c.image.shape[1], c.image.shape[0])
    assert len(lines) == len(sel) and len(sel) > 0
    for ln, s in zip(lines, sel):
        m = LINE.match(ln)
        assert m, ln
        assert int(m.group(1)) == 0
        assert m.group(2).endswith("mic000.mrc")
        assert float(m.group(3)) == pytest.approx(c.defocus, abs=1e-6)
        e3 = 360.0 - orc.psi_value(op.phi_step, int(s["psi_index"]))
        e3 = e3 - 360 if e3 >= 360 else e3
        assert float(m.group(6)) == pytest.approx(float(np.float32(eul[s["view"]][0])), abs=1e-5)
        assert float(m.group(7)) == pytest.approx(float(np.float32(eul[s["view"]][1])), abs=1e-5)
        assert float(m.group(8)) == pytest.approx(e3, abs=1e-6)
        assert float(m.group(9)) == s["cx"] and float(m.group(10)) == s["cy"]
        assert float(m.group(11)) == pytest.approx(s["score"], rel=1e-4, abs=2e-6)
