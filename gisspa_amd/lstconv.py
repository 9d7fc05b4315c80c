"""Pick-list formats on either side of the picker: the score filter and the
converters between GisSPA .lst files and RELION STAR files.

    python -m gisspa_amd.lstconv delete_bad_score <input1> <thres> <output>
    python -m gisspa_amd.lstconv lst2star <input lstfile> <scale factor> <pixel size> \\
                                          <out star file> <micrographs dir>
    python -m gisspa_amd.lstconv star2lst <relion STAR file> [--lst out.lst] [--allrelion 1]

Same arguments and output text as the reference's Python 2 scripts
(python_scripts/delete_bad_score.py:9-33,
python_scripts/convert_my-lst_to_relion_mic.py:9-53,
python_scripts/relion2lst_dont_write_mics.py:44-161).  Numbers they print
with str() follow Python 2's float formatting (12 significant digits, ".0"
on integral values), reproduced by py2_str.  Host-side text work: nothing
here touches the GPU.
"""
from __future__ import annotations

import math
import os
import sys

RELION_HEADER = (
    "# RELION; version 3.0-beta-2\n\ndata_\n\nloop_\n _rlnMicrographName #1 \n"
    "_rlnCoordinateX #2 \n_rlnCoordinateY #3 \n_rlnImageName #4 \n_rlnDefocusU #5 \n"
    "_rlnDefocusV #6 \n_rlnDefocusAngle #7 \n_rlnVoltage #8 \n_rlnSphericalAberration #9 \n"
    "_rlnAmplitudeContrast #10 \n_rlnMagnification #11 \n_rlnDetectorPixelSize #12 \n"
    "_rlnCtfFigureOfMerit #13 \n_rlnGroupNumber #14 \n_rlnAngleRot #15 \n_rlnAngleTilt #16 \n"
    "_rlnAnglePsi #17 \n_rlnOriginX #18 \n_rlnOriginY #19 \n_rlnClassNumber #20 \n"
    "_rlnNormCorrection #21 \n_rlnLogLikeliContribution #22 \n"
    "_rlnMaxValueProbDistribution #23 \n_rlnNrOfSignificantSamples #24\n")


def py2_str(v) -> str:
    """str() of a number as Python 2 prints it: floats with 12 significant
    digits ('%.12g') and '.0' appended when that reads as an integer."""
    if isinstance(v, bool) or isinstance(v, int):
        return str(v)
    v = float(v)
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    s = "%.12g" % v
    if not any(ch in s for ch in ".en"):
        s += ".0"
    return s


def _field_value(line: str, idx: int, k: int = 0) -> float:
    return float(line.split("\t")[idx].split("=")[1].split(",")[k])


# ---------------------------------------------------------------------------
# delete_bad_score.py:9-33
# ---------------------------------------------------------------------------
def delete_bad_score(lines: list[str], thres: float) -> list[str]:
    """Lines 4.. (the first three are skipped, :22) whose |score| exceeds
    thres; the score column is the field named 'score' in line 5 (:16-20)."""
    if len(lines) < 5:
        raise ValueError("delete_bad_score needs at least 5 lines (it reads line 5 for the "
                         "score column)")
    ref = lines[4]
    index = None
    for j in range(2, len(ref.split())):
        if ref.split("\t")[j].split("=")[0] == "score":
            index = j
    if index is None:
        raise ValueError("no score= field in line 5")
    return [ln for ln in lines[3:] if abs(float(ln.split("\t")[index].split("=")[1])) > thres]


# ---------------------------------------------------------------------------
# convert_my-lst_to_relion_mic.py:9-53
# ---------------------------------------------------------------------------
def lst_to_star(lines: list[str], scale: int, apix: float, mdir: str) -> str:
    """RELION 3.0 particle STAR text of a GisSPA pick list: micrograph
    mdir + basename, coordinates x scale, DefocusU/V = (defocus -/+ dfdiff)
    in A, rot = fmod(az - 90, 360), tilt = alt, psi = fmod(psi + 90, 360), the
    script's fixed 300 kV / 2.7 mm / 0.1 and magnification 50000 / apix."""
    out = [RELION_HEADER]
    mag = 50000 / apix
    for ln in lines:
        f = ln.split("\t")
        s1 = f[1]
        micro = mdir + s1.split("/")[-1]
        df = float(f[2].split("=")[1]) * 10000
        dfdiff = float(f[3].split("=")[1]) * 10000
        dfu, dfv = df - dfdiff, df + dfdiff
        dfang = float(f[4].split("=")[1])
        e1, e2, e3 = (_field_value(ln, 5, k) for k in range(3))
        cx = _field_value(ln, 6, 0) * scale
        cy = _field_value(ln, 6, 1) * scale
        out.append(
            micro + "\t" + py2_str(cx) + "\t" + py2_str(cy) + "\t000001@" + micro + "\t" +
            py2_str(dfu) + "\t" + py2_str(dfv) + "\t" + py2_str(dfang) +
            "\t300.000000\t2.700000\t0.100000\t" + py2_str(mag) +
            "\t5.000000\t0.140000\t1\t" + py2_str(math.fmod(e2 - 90, 360)) + "\t" + py2_str(e1) +
            "\t" + py2_str(math.fmod(e3 + 90, 360)) +
            "\t0.000000\t0.000000\t1\t0.600000\t1.800000e+05\t1.000000\t1" + "\n")
    return "".join(out)


# ---------------------------------------------------------------------------
# relion2lst_dont_write_mics.py:44-161 (STAR reading: EMAN2star.StarFile,
# not shipped; read here as one data block with a loop_ of _rln columns)
# ---------------------------------------------------------------------------
def _star_value(tok: str):
    for conv in (int, float):
        try:
            return conv(tok)
        except ValueError:
            pass
    return tok


def read_star(text: str) -> dict:
    """Columns of the first loop_ of a STAR file: {label without '_': values},
    values converted to int or float where they parse."""
    cols: list[str] = []
    rows: list[list] = []
    in_loop = False
    for raw in text.splitlines():
        line = raw.strip()
        if not in_loop:
            in_loop = line.startswith("loop_")
            continue
        if line.startswith("_") and not rows:
            cols.append(line.split()[0][1:])
            continue
        if not line or line.startswith("#"):
            continue
        if line.startswith("data_") or line.startswith("loop_"):
            break
        toks = line.split()
        if len(toks) < len(cols):
            raise ValueError(f"STAR row with {len(toks)} of {len(cols)} columns: {line}")
        rows.append([_star_value(t) for t in toks[:len(cols)]])
    return {c: [r[i] for r in rows] for i, c in enumerate(cols)}


def star_to_lst(star: dict, allrelion: bool = False) -> str:
    """GisSPA .lst text of a RELION particle STAR: per micrograph (runs of
    equal DefocusU, :114-120) the particle index restarts at 0; defocus =
    (U + V) / 2e4 um, dfdiff = |U - V| / 2e4, dfang + 90 (mod 360) when U > V
    (:129-137)."""
    names = star["rlnMicrographName"]
    du, dv, da = star["rlnDefocusU"], star["rlnDefocusV"], star["rlnDefocusAngle"]
    out = ["#LST\n"]
    fnum = 0
    keys = sorted(star.keys())
    for i in range(len(names)):
        if i == 0 or du[i - 1] != du[i]:
            fnum = 0
        dfu, dfv = du[i] / 1e4, dv[i] / 1e4
        dfang = da[i]
        defocus = (dfu + dfv) / 2.
        dfdiff = math.fabs(dfu - dfv) / 2.
        if dfu > dfv:
            dfang = math.fmod(dfang + 360. + 90., 360.)
        line = "%d\t%s\tdefocus=%s\tdfdiff=%s\tdfang=%s" % (
            fnum, names[i], py2_str(defocus), py2_str(dfdiff), py2_str(dfang))
        if allrelion:
            for var in keys:
                v = star[var][i]
                line += "\t%s=%s" % (var.split("rln")[-1],
                                     py2_str(v) if isinstance(v, (int, float)) else v)
        out.append(line + "\n")
        fnum += 1
    return "".join(out)


def main(argv: list[str]) -> int:
    if not argv:
        print(__doc__, file=sys.stderr)
        return 1
    tool, args = argv[0], argv[1:]
    if tool == "delete_bad_score":
        if len(args) < 3:
            print("<input1> <thres> <output>", file=sys.stderr)
            return 1
        with open(args[0]) as f:
            lines = f.readlines()
        with open(args[2], "w") as f:
            f.writelines(delete_bad_score(lines, float(args[1])))
        return 0
    if tool == "lst2star":
        if len(args) < 5:
            print("<input lstfile> <scale factor> <pixel size> <out star file> "
                  "<micrographs dir>", file=sys.stderr)
            return 1
        with open(args[0]) as f:
            lines = f.readlines()
        with open(args[3], "w") as f:
            f.write(lst_to_star(lines, int(args[1]), float(args[2]), args[4]))
        return 0
    if tool == "star2lst":
        import argparse
        ap = argparse.ArgumentParser(prog="star2lst")
        ap.add_argument("star")
        ap.add_argument("--lst", default="")
        ap.add_argument("--allrelion", type=int, default=0)
        a = ap.parse_args(args)
        with open(a.star) as f:
            star = read_star(f.read())
        lst = a.lst or os.path.basename(a.star)[:os.path.basename(a.star).rfind(".")] + ".lst"
        with open(lst, "w") as f:
            f.write(star_to_lst(star, bool(a.allrelion)))
        return 0
    print(f"unknown tool {tool!r}", file=sys.stderr)
    return 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
