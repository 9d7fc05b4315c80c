"""HDF5 image files (EMAN1 / EMAN2 layouts) through the C ABI's readers
(gs_image_count, gs_read_image, gs_hdf5_read, gs_hdf5_attr), host only.

What pins the parser:
  * a file written by libhdf5 itself: scipy ships a MATLAB 7.3 file
    (testhdf5_7.4_GLNX86.mat, an HDF5 file behind a 512-byte user block)
    holding the variable testdouble, and the same variable in a version-5
    MAT file that scipy.io.loadmat reads; the HDF5 reader must give those
    values, the dataset shape and the MATLAB_class string attribute;
  * the lookup3 checksum of the fixture writer against Bob Jenkins' published
    values;
  * files in both reference layouts (emdata_.cpp:174-205; EMAN1
    emhdf_.cpp:103-150, EMAN2 emhdf2_.cpp:245-363) from tests/_hdf5.py in
    both metadata styles (superblock 0 with symbol tables, superblock 2 with
    compact links), every layout (contiguous, compact, chunked with shuffle,
    deflate and per-chunk filter masks, single chunk), every element type;
    stacks must read back bit-exactly;
  * corrupt files (truncations, flipped bytes) end in an error, never a crash.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import gisspa_amd
from tests import _hdf5 as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _scipy_data(name):
    try:
        import scipy.io.matlab
    except ImportError:
        return None
    p = os.path.join(os.path.dirname(scipy.io.matlab.__file__), "tests", "data", name)
    return p if os.path.exists(p) else None


def test_file_written_by_libhdf5():
    h5 = _scipy_data("testhdf5_7.4_GLNX86.mat")
    v5 = _scipy_data("testdouble_7.1_GLNX86.mat")
    if not h5 or not v5:
        pytest.skip("scipy's MATLAB test files are not installed")
    import scipy.io
    ref = scipy.io.loadmat(v5)["testdouble"]          # 1 x 9 doubles
    got = gisspa_amd.hdf5_read(h5, "/testdouble")
    assert got.shape == (9, 1)                          # MATLAB stores column-major
    assert np.array_equal(got[:, 0], ref[0].astype(np.float32))
    assert gisspa_amd.hdf5_attr(h5, "/testdouble", "MATLAB_class") == "double"
    with pytest.raises(OSError, match="no object"):
        gisspa_amd.hdf5_read(h5, "/absent")
    with pytest.raises(OSError, match="no attribute"):
        gisspa_amd.hdf5_attr(h5, "/testdouble", "absent")


def test_lookup3_published_values():
    assert H.lookup3(b"") == 0xDEADBEEF
    assert H.lookup3(b"Four score and seven years ago") == 0x17770551
    assert H.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def _images(seed=0, n=7, shape=(40, 56)):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32) for _ in range(n)]


LAYOUTS = [dict(), dict(layout="compact"), dict(layout="chunked", chunks=(16, 24)),
           dict(layout="chunked", chunks=(16, 24), shuffle=True, deflate=6, raw_chunks=(1, 4)),
           dict(layout="chunked", chunks=(40, 56), deflate=4)]


@pytest.mark.parametrize("version", [0, 2])
@pytest.mark.parametrize("userblock", [0, 512])
@pytest.mark.parametrize("opts", LAYOUTS, ids=["contiguous", "compact", "chunked",
                                               "shuffle-deflate-masked", "single-chunk"])
def test_eman2_layout(tmp_path, version, userblock, opts):
    ims = _images()
    p = str(tmp_path / "stack.hdf")
    H.Writer(version, userblock).write(p, H.eman2_stack(ims, **opts))
    assert gisspa_amd.image_count(p) == len(ims)
    for i, im in enumerate(ims):
        assert np.array_equal(gisspa_amd.read_image(p, i), im)
    assert np.array_equal(gisspa_amd.read_image(p, -1), np.stack(ims))
    assert gisspa_amd.hdf5_attr(p, "/MDF/images/2", "EMAN.micrograph_id") == "synthetic"
    assert list(gisspa_amd.hdf5_attr(p, "/MDF/images/2", "EMAN.nx")) == [56.0]
    assert list(gisspa_amd.hdf5_attr(p, "/MDF/images", "imageid_max")) == [len(ims) - 1]
    with pytest.raises(OSError):
        gisspa_amd.read_image(p, len(ims))


@pytest.mark.parametrize("version", [0, 2])
def test_eman1_layout(tmp_path, version):
    """Root datasets "0", "1", ... with num_dataset; nx = dims[0]
    (emdata_.cpp:230-235)."""
    ims = _images(1, 5, (33, 47))
    p = str(tmp_path / "old.hdf")
    H.Writer(version).write(p, H.eman1_stack(ims))
    assert gisspa_amd.image_count(p) == 5
    for i, im in enumerate(ims):
        got = gisspa_amd.read_image(p, i)
        assert got.shape == (33, 47) and np.array_equal(got, im)
    # a negative index of a single image reads image 0 (emdata_.cpp:218)
    one = str(tmp_path / "one.hdf")
    H.Writer(version).write(one, H.eman1_stack(ims[:1]))
    assert np.array_equal(gisspa_amd.read_image(one, -1), ims[0])


@pytest.mark.parametrize("dtype", [">f4", "<f8", ">f8", "<f2", "<i2", ">i2", "|u1", "<i4",
                                   ">u2", "<i8"])
def test_element_types(tmp_path, dtype):
    rng = np.random.default_rng(3)
    dt = np.dtype(dtype)
    if dt.kind == "f":
        a = rng.standard_normal((24, 20)).astype(dt)
    else:
        info = np.iinfo(dt)
        lo, hi = max(info.min, -30000), min(info.max, 30000)
        a = rng.integers(lo, hi, (24, 20)).astype(dt)
    for version in (0, 2):
        p = str(tmp_path / f"t{version}.hdf")
        H.Writer(version).write(p, H.eman2_stack([a], layout="chunked", chunks=(7, 9),
                                                 shuffle=True, deflate=1))
        got = gisspa_amd.read_image(p, 0)
        assert np.array_equal(got, a.astype(np.float32)), dtype


def test_volume_and_mixed_sizes(tmp_path):
    rng = np.random.default_rng(4)
    vol = rng.standard_normal((6, 10, 12)).astype(np.float32)
    p = str(tmp_path / "vol.hdf")
    H.Writer(0).write(p, H.eman2_stack([vol]))
    assert np.array_equal(gisspa_amd.read_image(p, -1), vol)
    ims = [rng.standard_normal(s).astype(np.float32) for s in [(20, 20), (16, 16), (24, 24)]]
    q = str(tmp_path / "mixed.hdf")
    H.Writer(0).write(q, H.eman2_stack(ims))
    for i, im in enumerate(ims):
        assert np.array_equal(gisspa_amd.read_image(q, i), im)
    with pytest.raises(OSError, match="different sizes"):
        gisspa_amd.read_image(q, -1)


def test_mrc_through_the_same_entry(tmp_path):
    from gisspa_amd import mrc
    ims = np.stack(_images(5, 3, (18, 22)))
    p = str(tmp_path / "s.mrcs")
    mrc.write_mrc(p, ims)
    assert gisspa_amd.image_count(p) == 3
    assert np.array_equal(gisspa_amd.read_image(p, 1), ims[1])
    assert np.array_equal(gisspa_amd.read_image(p, -1), ims)


def test_not_an_image_file(tmp_path):
    p = tmp_path / "x.txt"
    p.write_text("not an image\n" * 100)
    with pytest.raises(OSError):
        gisspa_amd.read_image(str(p), 0)
    with pytest.raises(OSError):
        gisspa_amd.read_image(str(tmp_path / "missing.hdf"), 0)


CORRUPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import gisspa_amd
src = open(sys.argv[2], "rb").read()
rng = np.random.default_rng(int(sys.argv[3]))
bad = good = 0
for trial in range(300):
    b = bytearray(src)
    if trial % 3 == 0:
        b = b[:int(rng.integers(0, len(b)))]
    else:
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
    open(sys.argv[4], "wb").write(bytes(b))
    try:
        n = gisspa_amd.image_count(sys.argv[4])
        for i in range(min(n, 4)):
            gisspa_amd.read_image(sys.argv[4], i)
        good += 1
    except OSError:
        bad += 1
print("ok", good, bad)
"""


@pytest.mark.parametrize("version", [0, 2])
def test_corrupt_files_fail_cleanly(tmp_path, version):
    """Truncated files and flipped bytes (300 variants per style) end in
    OSError or in some decoded data, never in a crash of the process."""
    ims = _images(6, 4, (12, 16))
    p = str(tmp_path / "src.hdf")
    H.Writer(version).write(p, H.eman2_stack(ims, layout="chunked", chunks=(8, 8), shuffle=True,
                                             deflate=3))
    r = subprocess.run([sys.executable, "-c", CORRUPT, ROOT, p, str(version),
                        str(tmp_path / "bad.hdf")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("ok")


def test_cli_template_count_from_hdf(tmp_path):
    """./main counts the images of an HDF template stack like those of an MRC
    stack: fewer templates than Euler lines ends the run with a message."""
    from gisspa_amd import synth
    ims = _images(7, 3, (32, 32))
    H.Writer(0).write(str(tmp_path / "templates.hdf"), H.eman2_stack(ims))
    H.Writer(0).write(str(tmp_path / "mics.hdf"), H.eman2_stack([np.zeros((96, 96), np.float32)]))
    synth.write_euler_lst(str(tmp_path / "euler.lst"), synth.random_eulers(5, 1))
    synth.write_micrograph_lst(str(tmp_path / "m.lst"), [(0, "mics.hdf", 1.5, 0.0, 0.0)])
    cfg = tmp_path / "config"
    synth.write_config(str(cfg), input=str(tmp_path / "m.lst"),
                       template=str(tmp_path / "templates.hdf"),
                       eulerfile=str(tmp_path / "euler.lst"), output=str(tmp_path / "o.lst"),
                       first=0, last=1, angpix=1.0, phistep=30, kk=3, energy=300, cs=2.7,
                       Highres=8, Lowres=100, diameter=24, window_size=64, overlap=16)
    r = subprocess.run([gisspa_amd.MAIN_PATH, str(cfg)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "holds 3 templates, the Euler file lists 5" in r.stderr
    (tmp_path / "templates.hdf").write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(40))
    r = subprocess.run([gisspa_amd.MAIN_PATH, str(cfg)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Error" in r.stderr


def _mrc_bytes(data: np.ndarray, mode: int, big: bool) -> bytes:
    """An MRC file: 1024-byte header (nx, ny, nz, mode, ..., nlabl = 0) and
    the data, both in the given byte order."""
    nz, ny, nx = data.shape
    e = ">" if big else "<"
    hdr = np.zeros(256, dtype=e + "i4")
    hdr[0:4] = (nx, ny, nz, mode)
    dt = {0: "u1", 1: e + "i2", 2: e + "f4", 6: e + "u2"}[mode]
    return hdr.tobytes() + np.ascontiguousarray(data).astype(dt).tobytes()


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mrc_modes_and_byte_orders(tmp_path, mode, big):
    """readMRC's conversions (EMReader/emdata_.cpp:494-560): mode 0 bytes as
    b / 100.0 - 1.28 evaluated in double, mode 1 signed 16-bit, mode 2
    float; big-endian files detected from the header and swapped."""
    rng = np.random.default_rng(mode + 10 * big)
    if mode == 0:
        raw = rng.integers(0, 256, (3, 9, 11)).astype(np.uint8)
        want = (raw.astype(np.float64) / 100.0 - 1.28).astype(np.float32)
    elif mode == 1:
        raw = rng.integers(-32768, 32768, (3, 9, 11)).astype(np.int16)
        want = raw.astype(np.float32)
    else:
        raw = rng.standard_normal((3, 9, 11)).astype(np.float32)
        want = raw
    p = tmp_path / "x.mrc"
    p.write_bytes(_mrc_bytes(raw, mode, big))
    assert gisspa_amd.image_count(str(p)) == 3
    assert np.array_equal(gisspa_amd.read_image(str(p), -1), want)
    assert np.array_equal(gisspa_amd.read_image(str(p), 2), want[2])
    bad = tmp_path / "bad.mrc"
    bad.write_bytes(_mrc_bytes(raw, 6, big))   # modes above 4 are refused (emdata_.cpp:457)
    with pytest.raises(OSError):
        gisspa_amd.read_image(str(bad), 0)


def test_cyclic_group_btree_is_refused(tmp_path):
    """A group B-tree node that claims level 1 and names itself as its child
    (a cycle) is refused: children must sit exactly one level lower."""
    import struct
    ims = _images(8, 2, (8, 8))
    p = tmp_path / "c.hdf"
    H.Writer(0).write(str(p), H.eman1_stack(ims))
    b = bytearray(p.read_bytes())
    off = b.find(b"TREE")
    assert off > 0 and b[off + 4] == 0            # the root group's node
    b[off + 5] = 1                                # level 1 ...
    child = off + 8 + 16 + 8                      # after signature, siblings and key 0
    b[child:child + 8] = struct.pack("<Q", off)   # ... whose child is itself
    p.write_bytes(bytes(b))
    with pytest.raises(OSError, match="levels out of order"):
        gisspa_amd.image_count(str(p))


@pytest.mark.parametrize("suffix", [".mrc.gz", ".mrcs.bz2", ".mrc.bz"])
def test_compressed_mrc(tmp_path, suffix):
    """readMRC reads .bz / .bz2 files through bzcat (EMReader/emdata_.cpp:
    408-412); .gz files are read here too (zlib), the stack's sections
    included."""
    import bz2
    import gzip
    ims = np.stack(_images(9, 4, (20, 24)))
    raw = _mrc_bytes(ims, 2, False)
    p = tmp_path / ("s" + suffix)
    p.write_bytes(gzip.compress(raw) if suffix.endswith(".gz") else bz2.compress(raw))
    assert gisspa_amd.image_count(str(p)) == 4
    assert np.array_equal(gisspa_amd.read_image(str(p), -1), ims)
    assert np.array_equal(gisspa_amd.read_image(str(p), 3), ims[3])
    bad = tmp_path / ("t" + suffix)
    bad.write_bytes(b"not compressed data" * 10)
    with pytest.raises(OSError):   # zlib passes unknown bytes through: a short header
        gisspa_amd.read_image(str(bad), 0)
