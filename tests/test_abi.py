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
