"""Writes the golden fixtures of tests/test_golden.py: a seeded small picking
case (inputs from gisspa_amd.synth) with the FP64 oracle's per-map results,
pick list and the output lines the CLI prints for it (main.cu:985 format).

    python tests/golden/make_golden.py

The reference holds no test vectors for this path, so these fixtures are the
build's own: they pin the oracle and the generator against drift, and give
the GPU picker fixed inputs and expected outputs.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASE = dict(seed=0, l=128)


def make():
    from oracle import oracle as orc
    from tests import _cases
    c = _cases.tiny(**CASE)
    p = orc.params(c.l, **c.params)
    st = orc.pick_norm0(p, orc.ctf(c.defocus), c.templates, c.image)
    ny, nx = c.image.shape
    sel = orc.select_norm0(p, st, 0, nx, ny)
    eul = _euler(c)
    lines = []
    for r in sel:
        # printed psi: 360 - euler3 in float, wrapped (main.cu:1253-1254)
        psi = np.float32(360.0) - np.float32(orc.psi_value(c.params["phistep"], int(r["psi_index"])))
        if psi >= 360:
            psi = np.float32(psi - np.float32(360.0))
        lines.append("%d\t%s\tdefocus=%f\tdfdiff=%f\tdfang=%f\teuler=%f,%f,%f\tcenter=%f,%f\t"
                     "score=%f\n" % (0, "mic000.mrc", c.defocus, 0.0, 0.0,
                                     _as_read(eul[r["view"], 0]), _as_read(eul[r["view"], 1]),
                                     psi, np.float32(r["cx"]), np.float32(r["cy"]),
                                     np.float32(r["score"])))
    return c, st, sel, lines


def _as_read(v):
    """An Euler angle as the CLI holds it: written to the Euler file with six
    decimals (synth.write_euler_lst) and read back as a float."""
    return np.float32(float(f"{v:.6f}"))


def _euler(c):
    from gisspa_amd import synth
    return synth.random_eulers(c.templates.shape[0], CASE["seed"] + 12)


def main():
    c, st, sel, lines = make()
    np.savez_compressed(os.path.join(HERE, "tiny_seed0.npz"), templates=c.templates,
                        image=c.image, defocus=c.defocus, mapstats=st, picks=sel)
    with open(os.path.join(HERE, "tiny_seed0_picks.lst"), "w") as f:
        f.writelines(lines)
    print(f"{st.size} maps, {len(sel)} picks written")


if __name__ == "__main__":
    main()
