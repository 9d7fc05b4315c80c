"""Pick-list converters and the score filter (gisspa_amd/lstconv.py) on
hand-worked fixtures: Python 2's str() of floats, the score filter's skipped
header lines and |score| test, the lst -> STAR columns, and the STAR -> lst
defocus/astigmatism arithmetic with its per-micrograph particle index.
CPU only."""
import os
import subprocess
import sys

import pytest

from gisspa_amd import lstconv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("v,s", [(0.1, "0.1"), (1.0, "1.0"), (-0.0, "-0.0"), (1e16, "1e+16"),
                                 (1e11, "100000000000.0"), (1e12, "1e+12"),
                                 (1 / 3, "0.333333333333"), (2.5e-5, "2.5e-05"),
                                 (18000.000000000004, "18000.0"), (12345.678, "12345.678"),
                                 (7, "7")])
def test_python2_float_str(v, s):
    assert lstconv.py2_str(v) == s


def _line(i, score, path="m/a.mrc"):
    return (f"{i}\t{path}\tdefocus=1.800000\tdfdiff=0.100000\tdfang=30.000000\t"
            f"euler=10.000000,20.000000,30.000000\tcenter=100.500000,50.250000\t"
            f"score={score:f}\n")


def test_delete_bad_score():
    lines = [_line(0, 99), _line(1, 99), _line(2, 99), _line(3, 7.5), _line(4, 8.0),
             _line(5, -9.0), _line(6, 8.000001)]
    out = lstconv.delete_bad_score(lines, 8.0)
    # lines 0-2 are never written; |score| > 8 strictly
    assert out == [lines[5], lines[6]]
    with pytest.raises(ValueError):
        lstconv.delete_bad_score(lines[:4], 8.0)


def test_lst_to_star_columns():
    txt = lstconv.lst_to_star([_line(0, 9.5)], 2, 1.0, "mics/")
    assert txt.startswith(lstconv.RELION_HEADER)
    row = txt[len(lstconv.RELION_HEADER):].rstrip("\n").split("\t")
    # micrograph, x, y, image, U, V, angle, 300 kV, Cs, Q0, mag, pixel, FOM, group,
    # rot = az - 90, tilt = alt, psi = psi + 90, ...
    assert row[0] == "mics/a.mrc" and row[3] == "000001@mics/a.mrc"
    assert row[1:3] == ["201.0", "100.5"]
    assert row[4:7] == ["17000.0", "19000.0", "30.0"]
    assert row[10] == "50000.0"
    assert row[14:17] == ["-70.0", "10.0", "120.0"]
    assert len(row) == 24


STAR = """
data_

loop_
_rlnMicrographName #1
_rlnDefocusU #2
_rlnDefocusV #3
_rlnDefocusAngle #4
_rlnCoordinateX #5
m/one.mrc 18000 17000 45 10
m/one.mrc 18000 17000 45 20
m/two.mrc 25000.5 25000.5 12.5 30
"""


def test_star_to_lst():
    star = lstconv.read_star(STAR)
    assert star["rlnDefocusU"] == [18000, 18000, 25000.5]
    out = lstconv.star_to_lst(star).splitlines()
    assert out[0] == "#LST"
    assert out[1] == "0\tm/one.mrc\tdefocus=1.75\tdfdiff=0.05\tdfang=135.0"
    assert out[2] == "1\tm/one.mrc\tdefocus=1.75\tdfdiff=0.05\tdfang=135.0"
    # a new DefocusU starts a new micrograph: index back to 0, U == V keeps the angle
    assert out[3] == "0\tm/two.mrc\tdefocus=2.50005\tdfdiff=0.0\tdfang=12.5"
    full = lstconv.star_to_lst(star, allrelion=True).splitlines()
    assert full[1].endswith("\tCoordinateX=10\tDefocusAngle=45\tDefocusU=18000\t"
                            "DefocusV=17000\tMicrographName=m/one.mrc")


def test_cli(tmp_path):
    src = tmp_path / "picks.lst"
    src.write_text("".join(_line(i, 5 + i) for i in range(8)))
    out = tmp_path / "good.lst"
    r = subprocess.run([sys.executable, "-m", "gisspa_amd.lstconv", "delete_bad_score",
                        str(src), "9", str(out)], capture_output=True, text=True, cwd=ROOT,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == "".join(_line(i, 5 + i) for i in range(5, 8))
    star = tmp_path / "p.star"
    star.write_text(STAR)
    r = subprocess.run([sys.executable, "-m", "gisspa_amd.lstconv", "star2lst", str(star),
                        "--lst", str(tmp_path / "p.lst")], capture_output=True, text=True,
                       cwd=ROOT, timeout=60)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "p.lst").read_text().startswith("#LST\n0\tm/one.mrc")
