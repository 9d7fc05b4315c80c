"""Minimal MRC (mode 2, float32) stack reader/writer used by the synthetic
data generator and the tests.  The 1024-byte header follows the layout the
reference reads (EMReader/ImgType.h:23-83, EMReader/emdata_.cpp:384-617)."""
from __future__ import annotations

import struct

import numpy as np


def write_mrc(path: str, data: np.ndarray) -> None:
    a = np.asarray(data, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    hdr = bytearray(1024)
    struct.pack_into("<3i", hdr, 0, nx, ny, nz)
    struct.pack_into("<i", hdr, 12, 2)                 # mode 2 = float
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)       # mx my mz
    struct.pack_into("<3f", hdr, 40, float(nx), float(ny), float(nz))
    struct.pack_into("<3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)
    struct.pack_into("<3f", hdr, 76, float(a.min()), float(a.max()), float(a.mean()))
    hdr[208:212] = b"MAP "
    struct.pack_into("<i", hdr, 212, 0x00004444)
    struct.pack_into("<i", hdr, 220, 0)                # nlabl
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.ascontiguousarray(a).astype("<f4").tobytes())


def read_mrc(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        hdr = f.read(1024)
        nx, ny, nz, mode = struct.unpack_from("<4i", hdr, 0)
        if mode != 2:
            raise ValueError(f"{path}: only mode 2 supported here, got {mode}")
        nsym = struct.unpack_from("<i", hdr, 92)[0]
        f.seek(1024 + nsym)
        a = np.frombuffer(f.read(4 * nx * ny * nz), dtype="<f4").reshape(nz, ny, nx)
    return a.copy()
