"""C ABI surface (CPU): the library loads, exports every entry point the
header declares, and fails cleanly (error codes, no exit) without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gisspa_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = gisspa_amd.load()
    names = gisspa_amd.exported_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_params_defaults_match_reference():
    p = gisspa_amd.make_params()
    # EMReader/DataReader.h:16-42
    assert p.threshold == pytest.approx(7.0)
    assert p.window_size == 320 and p.overlap == -1 and p.phase_flip == 0
    assert p.norm_type == 0 and p.invert == 0
    assert p.ampconst == pytest.approx(0.07)
    assert p.a == pytest.approx(-9.32) and p.bfactor3 == pytest.approx(1.28732)
    assert p.apix == -1 and p.diameter == -1


def test_errors_are_codes_not_exits():
    lib = gisspa_amd.load()
    for code in range(-6, 1):
        assert lib.gs_strerror(code)
    p = gisspa_amd.make_params()   # required keys unset
    ctx = ctypes.c_void_p()
    assert lib.gs_create(ctypes.byref(p), 0, ctypes.byref(ctx)) == -1
    p = gisspa_amd.make_params(angpix=1, phistep=5, kk=3, energy=300, cs=2.7, Highres=8,
                               Lowres=100, diameter=100)
    r = lib.gs_create(ctypes.byref(p), 99, ctypes.byref(ctx))
    assert r in (-3,)  # no such device here (or anywhere)
    assert lib.gs_pick(None, 0, -1, None, 0, ctypes.byref(ctypes.c_size_t())) == -1


def test_header_declares_only_c_types():
    hdr = open(os.path.join(ROOT, "include", "gisspa.h")).read()
    assert "extern \"C\"" in hdr
    for bad in ("torch", "std::", "hipStream_t", "float2"):
        assert bad not in hdr


def test_cli_help_and_missing_keys(tmp_path):
    main = gisspa_amd.MAIN_PATH
    r = subprocess.run([main, "-h"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "./main config_file" in r.stdout
    cfg = tmp_path / "config"
    cfg.write_text("# only a few keys\nangpix = 1.0\nphistep 5\nbogus_key = 3\n")
    r = subprocess.run([main, str(cfg)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1
    assert "Undefined option: 3 .Abort." in r.stdout        # DataReader.cpp:245 wording
    assert "Error : template file is requested." in r.stdout
    assert "--angpix           1.000000" in r.stdout
    assert "--phistep          5.000000" in r.stdout
    assert "--micrograph_batch 4" in r.stdout
    cfg.write_text("angpix = 1.0\nmicrograph_batch = 0\n")
    r = subprocess.run([main, str(cfg)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "micrograph_batch must be at least 1." in r.stdout


def test_remove_repeats_entry_and_tool_errors(tmp_path):
    lib = gisspa_amd.load()
    nk = ctypes.c_size_t(7)
    assert lib.gs_remove_repeats(0, None, 0, 10.0, 5.0, None, ctypes.byref(nk)) == 0
    assert nk.value == 0
    recs = gisspa_amd.pick_records([0], [1.0], [2.0], [0.0], [0.0], [0.0], [8.0])
    keep = np.zeros(1, np.uint8)
    assert lib.gs_remove_repeats(0, None, 1, 10.0, 5.0, keep.ctypes.data, None) == -1
    if not gpu_here():
        assert lib.gs_remove_repeats(0, recs.ctypes.data, 1, 10.0, 5.0, keep.ctypes.data,
                                     None) == -3
    tool = os.path.join(os.path.dirname(gisspa_amd.LIB_PATH), "remove_repeats")
    run = lambda *a: subprocess.run([tool, *map(str, a)], capture_output=True, text=True,
                                    timeout=30)
    assert run().returncode == 1
    empty = tmp_path / "empty.lst"
    empty.write_text("")
    assert run(empty, 1, 10, 5, tmp_path / "o").returncode == 1
    bad = tmp_path / "bad.lst"
    bad.write_text("0\ta.mrc\tdefocus=1\n")
    r = run(bad, 1, 10, 5, tmp_path / "o")
    assert r.returncode == 1 and "not a pick line" in r.stderr
    line = ("0\t{}\tdefocus=1.0\tdfdiff=0\tdfang=0\teuler=1,2,3\tcenter=10,20\t"
            "score=8.5\n")
    two = tmp_path / "two.lst"
    two.write_text(line.format("a.mrc") + line.format("b.mrc"))
    r = run(two, 1, 10, 5, tmp_path / "o")
    assert r.returncode == 1 and "exceed" in r.stderr


def gpu_here() -> bool:
    from tests.conftest import gpu_available
    return gpu_available()


def test_project_entry_and_tool_errors(tmp_path):
    lib = gisspa_amd.load()
    vol = np.zeros((4, 4, 4), np.float32)
    eul = np.zeros(3, np.float32)
    out = np.zeros((1, 4, 4), np.float32)
    assert lib.gs_project(0, None, 4, eul.ctypes.data, 1, out.ctypes.data) == -1
    assert lib.gs_project(0, vol.ctypes.data, 1, eul.ctypes.data, 1, out.ctypes.data) == -1
    assert lib.gs_project(0, vol.ctypes.data, 4, eul.ctypes.data, 0, out.ctypes.data) == 0
    if not gpu_here():
        assert lib.gs_project(0, vol.ctypes.data, 4, eul.ctypes.data, 1, out.ctypes.data) == -3
    from gisspa_amd import mrc
    tool = os.path.join(os.path.dirname(gisspa_amd.LIB_PATH), "project3d")
    assert subprocess.run([tool], capture_output=True, timeout=30).returncode == 1
    mrc.write_mrc(str(tmp_path / "flat.mrc"), np.zeros((3, 4, 4), np.float32))
    (tmp_path / "e.lst").write_text("0 0 0\n")
    r = subprocess.run([tool, str(tmp_path / "flat.mrc"), str(tmp_path / "e.lst"),
                        str(tmp_path / "o.mrcs")], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "cubic" in r.stderr
