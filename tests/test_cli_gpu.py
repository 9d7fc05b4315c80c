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


def test_cli_micrograph_batches(gpu, tmp_path):
    """Five micrographs of different sizes and defoci: searched one by one
    (micrograph_batch = 1, the reference's loop), in groups of two (a ragged
    last group) and all five together, the output files are identical (up
    to the folder in the micrograph paths)
    and the "finished" messages come in micrograph order.  The one-by-one
    run's picks equal the oracle's, micrograph by micrograph (templates
    weighted by the first micrograph's CTF, main.cu:1106)."""
    sizes = [(320, 288), (256, 352), (400, 200), (288, 288), (352, 256)]
    defs = [1.5, 1.2, 1.9, 1.6, 2.1]
    cases = [_cases.tiny(seed=30 + i, l=96, size=sz, threshold=4.0) for i, sz in enumerate(sizes)]
    c0 = cases[0]
    eul = synth.random_eulers(len(c0.templates), 9)
    mics = [(c.image, d) for c, d in zip(cases, defs)]
    outs = {}
    for group in (1, 2, 5):
        d = tmp_path / f"g{group}"
        wl = synth.Workload("cli", c0.templates, eul, mics, dict(c0.params), [])
        cfg = synth.write_workload(wl, str(d), extra=dict(micrograph_batch=group))
        r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
        print(r.stdout[-1000:], r.stderr[-1000:])
        assert r.returncode == 0
        done = [int(m) for m in re.findall(r"Image (\d+) finished", r.stdout)]
        assert done == list(range(5))
        # lines name the micrographs by path: compare with the folder masked
        outs[group] = open(os.path.join(d, "picks.lst")).read().replace(str(d), "DIR")
    assert outs[1] == outs[2] == outs[5]
    lines = outs[1].splitlines()
    op = orc.params(c0.l, **c0.params)
    treal = orc.template_prep(op, orc.ctf(defs[0]), c0.templates)
    P = orc.num_psi(op.phi_step)
    k = 0
    for i, c in enumerate(cases):
        tiles = orc.micrograph_prep(op, orc.ctf(defs[i]), c.image)
        st = np.stack([orc.correlate_many(op, tiles, orc.rotate_normalize(
            op, treal, orc.psi_value(op.phi_step, ip))) for ip in range(P)])
        sel = orc.select_norm0(op, st, 0, c.image.shape[1], c.image.shape[0])
        mine = lines[k:k + len(sel)]
        k += len(sel)
        for ln, s in zip(mine, sel):
            m = LINE.match(ln)
            assert m, ln
            assert int(m.group(1)) == i and m.group(2).endswith(f"mic{i:03d}.mrc")
            assert float(m.group(9)) == s["cx"] and float(m.group(10)) == s["cy"]
            assert float(m.group(11)) == pytest.approx(s["score"], rel=1e-4, abs=2e-6)
    assert k == len(lines)


def _lines_without_paths(path):
    out = []
    for ln in open(path).read().splitlines():
        f = ln.split("\t")
        out.append("\t".join([f[0]] + f[2:]))
    return out


def test_cli_reads_hdf_files(gpu, tmp_path):
    """The same workload with its templates and micrographs in HDF files
    (emdata::readImage, EMReader/emdata_.cpp:174-205) gives the lines of the
    MRC run.  Run A: EMAN2 layout, templates with version-2 metadata and one
    deflated single-chunk dataset per image, the two micrographs (of
    different sizes) as images 0 and 1 of one chunked, shuffled and deflated
    stack.  Run B: EMAN1 layout, templates behind a 512-byte user block."""
    from tests import _hdf5 as H
    c = _cases.tiny(seed=2, l=128, n_views=5, phistep=45.0, threshold=4.5)
    c2 = _cases.tiny(seed=3, l=128, n_views=5, phistep=45.0, threshold=4.5, size=(288, 320))
    eul = synth.random_eulers(len(c.templates), 7)
    mics = [(c.image, c.defocus), (c2.image, 1.7)]
    wl = synth.Workload("cli", c.templates, eul, mics, dict(c.params), [])
    d = tmp_path / "mrc"
    cfg = synth.write_workload(wl, str(d))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = _lines_without_paths(os.path.join(d, "picks.lst"))
    assert len(ref) > 0
    tm = c.templates.shape[1]
    runs = {
        "A": (lambda p: H.Writer(2).write(p, H.eman2_stack(list(c.templates), layout="chunked",
                                                           chunks=(tm, tm), deflate=4)),
              lambda p: H.Writer(0).write(p, H.eman2_stack([c.image, c2.image], layout="chunked",
                                                           chunks=(64, 96), shuffle=True,
                                                           deflate=6))),
        "B": (lambda p: H.Writer(0, 512).write(p, H.eman1_stack(list(c.templates))),
              lambda p: H.Writer(0).write(p, H.eman1_stack([c.image, c2.image]))),
    }
    for name, (write_templates, write_mics) in runs.items():
        h = tmp_path / name
        cfg = synth.write_workload(wl, str(h))
        write_templates(str(h / "templates.hdf"))
        write_mics(str(h / "micrographs.hdf"))
        synth.write_micrograph_lst(str(h / "micrographs.lst"),
                                   [(i, "micrographs.hdf", df, 0.0, 0.0) for i, (_, df) in
                                    enumerate(mics)])
        text = open(cfg).read().replace("templates.mrcs", "templates.hdf")
        open(cfg, "w").write(text)
        r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.stderr[-2000:])
        got = open(os.path.join(h, "picks.lst")).read().splitlines()
        assert all(ln.split("\t")[1].endswith("micrographs.hdf") for ln in got)
        assert _lines_without_paths(os.path.join(h, "picks.lst")) == ref, name


def test_cli_hdf_templates_of_mixed_boxes(gpu, tmp_path):
    """An EMAN2 HDF template stack whose images have their own sizes (64,
    48 and 57 px): the CLI centres each in the window by its size
    (main.cu:182-223) and prints the picks the library gives for the same
    templates (gs_set_templates_mixed)."""
    import gisspa_amd
    from tests import _hdf5 as H
    c = _cases.tiny(seed=6, l=128, box=64, n_views=6, threshold=4.0)
    boxes = [64, 48, 57, 64, 48, 57]
    tms = []
    for t, b in zip(c.templates, boxes):
        o = (64 - b) // 2
        tms.append(np.ascontiguousarray(t[o:o + b, o:o + b]))
    eul = synth.random_eulers(len(tms), 5)
    wl = synth.Workload("cli", c.templates, eul, [(c.image, c.defocus)], dict(c.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    H.Writer(0).write(str(tmp_path / "templates.hdf"), H.eman2_stack(tms))
    text = open(cfg).read().replace("templates.mrcs", "templates.hdf")
    open(cfg, "w").write(text)
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = open(os.path.join(tmp_path, "picks.lst")).read().splitlines()
    with gisspa_amd.Picker(0, **c.params) as pk:
        pk.set_templates_mixed(tms, c.defocus)
        pk.set_micrograph(c.image, c.defocus)
        picks = pk.pick()
    assert len(lines) == len(picks) and len(picks) > 0
    for ln, p in zip(lines, picks):
        m = LINE.match(ln)
        assert m, ln
        assert float(m.group(6)) == pytest.approx(float(np.float32(eul[p["view"]][0])), abs=1e-5)
        assert float(m.group(9)) == p["cx"] and float(m.group(10)) == p["cy"]
        assert float(m.group(11)) == pytest.approx(float(p["score"]), rel=1e-5, abs=2e-6)


def test_cli_norm_type1(gpu, tmp_path):
    """norm_type = 1 through ./main (main.cu:1170-1229): the output lines
    hold the oracle's run records (z-scores within 1e-4; a record within
    1e-4 of the threshold may be missing on either side)."""
    c = _cases.tiny(seed=3, l=64, box=48, n_views=4, size=(200, 176), phistep=45.0, overlap=16,
                    threshold=3.0)
    c.params["norm_type"] = 1
    eul = synth.random_eulers(len(c.templates), 13)
    wl = synth.Workload("cli", c.templates, eul, [(c.image, c.defocus)], dict(c.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = open(os.path.join(tmp_path, "picks.lst")).read().splitlines()
    op = orc.params(c.l, **c.params)
    ref = orc.pick_norm1(op, orc.ctf(c.defocus), c.templates, c.image)
    thr = c.params["threshold"]
    rk = {}
    for s in ref:
        e3 = 360.0 - orc.psi_value(op.phi_step, int(s["psi_index"]))
        e3 = e3 - 360 if e3 >= 360 else e3
        rk[(round(e3, 4), int(s["view"]), float(s["cx"]), float(s["cy"]))] = float(s["score"])
    gk = {}
    for ln in lines:
        m = LINE.match(ln)
        assert m, ln
        view = [j for j in range(len(eul))
                if abs(float(m.group(6)) - float(np.float32(eul[j][0]))) < 1e-5 and
                abs(float(m.group(7)) - float(np.float32(eul[j][1]))) < 1e-5]
        assert len(view) == 1, ln
        gk[(round(float(m.group(8)), 4), view[0], float(m.group(9)), float(m.group(10)))] = \
            float(m.group(11))
    print("norm1 CLI lines", len(gk), "oracle", len(rk))
    assert len(rk) > 10
    for k in set(rk) ^ set(gk):
        s = rk.get(k, gk.get(k))
        assert abs(s - thr) <= 1e-4 * thr + 2e-6, (k, s)
    for k in set(rk) & set(gk):
        assert abs(rk[k] - gk[k]) <= 1e-4 * max(1.0, abs(rk[k])) + 2e-6, (k, rk[k], gk[k])


def _write_lsx(path, entries, width=96):
    """#LSX list (fixed-width records after a header giving the record size,
    DataReader.cpp:330-347)."""
    with open(path, "w") as f:
        f.write("#LSX\n# fixed-width records: every line below has the size given next\n")
        f.write(f"# {width}\n")
        for idx, p, df, dfdiff, dfang in entries:
            rec = f"{idx}\t{p}\tdefocus={df:.6f}\tdfdiff={dfdiff:.6f}\tdfang={dfang:.6f}"
            assert len(rec) < width
            f.write(rec.ljust(width - 1) + "\n")


def test_cli_lsx_list(gpu, tmp_path):
    """The same three micrographs listed as #LSX (fixed-width records read
    by seeking) and as #LST give the same output file."""
    sizes = [(288, 256), (256, 320), (320, 288)]
    cases = [_cases.tiny(seed=40 + i, l=96, size=sz, threshold=4.0) for i, sz in enumerate(sizes)]
    c0 = cases[0]
    eul = synth.random_eulers(len(c0.templates), 21)
    mics = [(c.image, 1.4 + 0.2 * i) for i, c in enumerate(cases)]
    wl = synth.Workload("cli", c0.templates, eul, mics, dict(c0.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lst_out = open(os.path.join(tmp_path, "picks.lst")).read()
    assert lst_out.count("\n") > 0
    _write_lsx(str(tmp_path / "micrographs.lsx"),
               [(i, f"mic{i:03d}.mrc", df, 0.0, 0.0) for i, (_, df) in enumerate(mics)])
    text = open(cfg).read().replace("micrographs.lst", "micrographs.lsx").replace(
        "picks.lst", "picks_lsx.lst")
    open(cfg, "w").write(text)
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert open(os.path.join(tmp_path, "picks_lsx.lst")).read() == lst_out


def test_cli_reads_the_reference_euler_list(gpu, tmp_path):
    """The Euler file in the reference's own text layout (right-aligned
    8-column values split by tabs, as Euler_file/euler_C1_step5.lst, each
    value the shortest text that reads back as the same float): twelve of
    its lines, with
    templates projected at those views; every output line carries the
    file's alt and az (as parsed by sscanf %f, DataReader.cpp:267-284) and
    the picks are the oracle's."""
    views = _cases.c1_step5_views()[::137][:12]
    c = _cases.tiny(seed=4, l=128, n_views=12, phistep=60.0, threshold=4.0)
    dens = synth.blob_density(20, 64, 15, sigma_range=(1.5, 4.0))
    tm = synth.project_templates(dens, 64, views)
    wl = synth.Workload("cli-euler", tm, views, [(c.image, c.defocus)], dict(c.params), [])
    cfg = synth.write_workload(wl, str(tmp_path))
    short = lambda v: np.format_float_positional(np.float32(v), trim="-")
    with open(os.path.join(tmp_path, "euler.lst"), "w") as f:
        for alt, az in views:
            f.write("%8s\t%8s\t%8s\n" % (short(alt), short(az), "0"))
    r = subprocess.run([MAIN_PATH, cfg], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    lines = open(os.path.join(tmp_path, "picks.lst")).read().splitlines()
    op = orc.params(c.l, **c.params)
    st = orc.pick_norm0(op, orc.ctf(c.defocus), tm, c.image)
    sel = orc.select_norm0(op, st, 0, c.image.shape[1], c.image.shape[0])
    assert len(lines) == len(sel) > 0
    for ln, s in zip(lines, sel):
        m = LINE.match(ln)
        assert m, ln
        alt, az = views[s["view"]]
        assert float(m.group(6)) == pytest.approx(float(np.float32(alt)), abs=1e-5)
        assert float(m.group(7)) == pytest.approx(float(np.float32(az)), abs=1e-5)
        assert float(m.group(9)) == s["cx"] and float(m.group(10)) == s["cy"]
        assert float(m.group(11)) == pytest.approx(s["score"], rel=1e-4, abs=2e-6)
