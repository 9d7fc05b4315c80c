"""Minimal HDF5 writer for test fixtures (no h5py or libhdf5 in this image).

Writes the structures of the HDF5 file format specification (version 3.0)
that EMAN's image files use, in two styles:
  * version=0: superblock 0, version-1 object headers, groups as symbol tables
    (v1 B-tree + local heap + symbol table nodes), version-1 dataspace and
    attribute messages -- what libhdf5 1.8 writes by default;
  * version=2: superblock 2, version-2 object headers ("OHDR") with lookup3
    checksums, groups as compact link messages, version-2 dataspaces,
    version-3 attributes, layout version 4 for single-chunk datasets.
Datasets are contiguous, compact or chunked (v1 B-tree index, or a single
chunk in the version-2 style) with optional shuffle / deflate filters and
per-chunk filter masks.  Test infrastructure only: the picker never writes
HDF5.
"""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIG = b"\x89HDF\r\n\x1a\n"
M32 = 0xFFFFFFFF


def lookup3(data: bytes, initval: int = 0) -> int:
    """Bob Jenkins' lookup3 hashlittle, the checksum of version-2 HDF5
    metadata (byte-wise tail, as the format's reference implementation)."""
    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & M32

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M32
    i = 0
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], "little")) & M32
        b = (b + int.from_bytes(data[i + 4:i + 8], "little")) & M32
        c = (c + int.from_bytes(data[i + 8:i + 12], "little")) & M32
        a = (a - c) & M32; a ^= rot(c, 4); c = (c + b) & M32
        b = (b - a) & M32; b ^= rot(a, 6); a = (a + c) & M32
        c = (c - b) & M32; c ^= rot(b, 8); b = (b + a) & M32
        a = (a - c) & M32; a ^= rot(c, 16); c = (c + b) & M32
        b = (b - a) & M32; b ^= rot(a, 19); a = (a + c) & M32
        c = (c - b) & M32; c ^= rot(b, 4); b = (b + a) & M32
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = bytes(data[i:]) + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & M32
    b = (b + int.from_bytes(tail[4:8], "little")) & M32
    c = (c + int.from_bytes(tail[8:12], "little")) & M32
    c ^= b; c = (c - rot(b, 14)) & M32
    a ^= c; a = (a - rot(c, 11)) & M32
    b ^= a; b = (b - rot(a, 25)) & M32
    c ^= b; c = (c - rot(b, 16)) & M32
    a ^= c; a = (a - rot(c, 4)) & M32
    b ^= a; b = (b - rot(a, 14)) & M32
    c ^= b; c = (c - rot(b, 24)) & M32
    return c


def _pad8(b: bytes) -> bytes:
    return b + bytes((-len(b)) % 8)


@dataclass
class Dataset:
    data: np.ndarray
    attrs: dict = field(default_factory=dict)
    layout: str = "contiguous"          # contiguous | compact | chunked
    chunks: tuple | None = None
    shuffle: bool = False
    deflate: int | None = None          # zlib level
    raw_chunks: tuple = ()              # chunk indices stored with the deflate step skipped


@dataclass
class Group:
    members: dict = field(default_factory=dict)
    attrs: dict = field(default_factory=dict)


def _dtype_msg(dt: np.dtype) -> bytes:
    dt = np.dtype(dt)
    big = 1 if dt.byteorder == ">" else 0
    if dt.kind == "f":
        size = dt.itemsize
        prec, eloc, esz, msz, bias = {2: (16, 10, 5, 10, 15), 4: (32, 23, 8, 23, 127),
                                      8: (64, 52, 11, 52, 1023)}[size]
        bits = bytes([0x20 | big, prec - 1, 0])
        return (bytes([0x11]) + bits + struct.pack("<I", size) +
                struct.pack("<HHBBBBI", 0, prec, eloc, esz, 0, msz, bias))
    if dt.kind in "iu":
        bits = bytes([big | (8 if dt.kind == "i" else 0), 0, 0])
        return (bytes([0x10]) + bits + struct.pack("<I", dt.itemsize) +
                struct.pack("<HH", 0, 8 * dt.itemsize))
    raise ValueError(f"unsupported dtype {dt}")


def _string_type(n: int) -> bytes:
    return bytes([0x13, 0, 0, 0]) + struct.pack("<I", n)


class Writer:
    def __init__(self, version: int = 0, userblock: int = 0):
        assert version in (0, 2)
        assert userblock in (0, 512, 1024)
        self.v = version
        self.userblock = userblock
        self.buf = bytearray()

    # -- raw space -----------------------------------------------------------
    def _alloc(self, blob: bytes) -> int:
        while len(self.buf) % 8:
            self.buf.append(0)
        addr = len(self.buf)
        self.buf += blob
        return addr

    # -- messages --------------------------------------------------------------
    def _space(self, shape) -> bytes:
        rank = len(shape)
        if self.v == 0:
            return struct.pack("<BBB5x", 1, rank, 0) + b"".join(struct.pack("<Q", d) for d in shape)
        return struct.pack("<BBBB", 2, rank, 0, 1 if rank else 0) + b"".join(
            struct.pack("<Q", d) for d in shape)

    def _attr(self, name: str, value) -> bytes:
        if isinstance(value, str):
            raw = value.encode() + b"\0"
            dt, shape = _string_type(len(raw)), ()
        else:
            arr = np.asarray(value)
            if arr.dtype == np.int64 and arr.ndim == 0:
                arr = arr.astype(np.int32)
            dt, shape, raw = _dtype_msg(arr.dtype), arr.shape, arr.tobytes()
        sp = self._space(shape)
        nm = name.encode() + b"\0"
        if self.v == 0:
            return (struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(sp)) + _pad8(nm) +
                    _pad8(dt) + _pad8(sp) + raw)
        return struct.pack("<BBHHHB", 3, 0, len(nm), len(dt), len(sp), 0) + nm + dt + sp + raw

    def _ohdr(self, msgs) -> int:
        if self.v == 0:
            body = b""
            for t, d in msgs:
                d = _pad8(d)
                body += struct.pack("<HHB3x", t, len(d), 0) + d
            return self._alloc(struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + bytes(4) +
                               body)
        body = b"".join(struct.pack("<BHB", t, len(d), 0) + d for t, d in msgs)
        blob = b"OHDR" + bytes([2, 0x02]) + struct.pack("<I", len(body)) + body
        return self._alloc(blob + struct.pack("<I", lookup3(blob)))

    # -- objects -----------------------------------------------------------------
    def _dataset(self, ds: Dataset) -> int:
        arr = np.asarray(ds.data)
        msgs = [(1, self._space(arr.shape)), (3, _dtype_msg(arr.dtype))]
        es = arr.dtype.itemsize
        filters = []
        if ds.shuffle:
            filters.append((2, [es]))
        if ds.deflate is not None:
            filters.append((1, [ds.deflate]))
        if ds.layout == "contiguous":
            addr = self._alloc(arr.tobytes())
            layout = struct.pack("<BBQQ", 3, 1, addr, arr.nbytes)
        elif ds.layout == "compact":
            layout = struct.pack("<BBH", 3, 0, arr.nbytes) + arr.tobytes()
        else:
            cd = tuple(ds.chunks)
            grid = [-(-n // c) for n, c in zip(arr.shape, cd)]
            recs = []
            for k, idx in enumerate(np.ndindex(*grid)):
                off = [i * c for i, c in zip(idx, cd)]
                chunk = np.zeros(cd, arr.dtype)
                sl = tuple(slice(o, min(o + c, n)) for o, c, n in zip(off, cd, arr.shape))
                chunk[tuple(slice(0, s.stop - s.start) for s in sl)] = arr[sl]
                raw = chunk.tobytes()
                mask = 0
                for fi, (fid, cdv) in enumerate(filters):
                    if fid == 2:
                        raw = np.frombuffer(raw, np.uint8).reshape(-1, es).T.tobytes()
                    elif k in ds.raw_chunks:
                        mask |= 1 << fi
                    else:
                        raw = zlib.compress(raw, cdv[0])
                recs.append((self._alloc(raw), len(raw), mask, off))
            if self.v == 2 and len(recs) == 1 and tuple(cd) == arr.shape:
                a, n, mask, _ = recs[0]
                flags = 2 if filters else 0
                layout = struct.pack("<BBBBB", 4, 2, flags, arr.ndim + 1, 4) + b"".join(
                    struct.pack("<I", c) for c in cd + (es,)) + bytes([1])
                if filters:
                    layout += struct.pack("<QI", n, mask)
                layout += struct.pack("<Q", a)
            else:
                node = b"TREE" + struct.pack("<BBHQQ", 1, 0, len(recs), UNDEF, UNDEF)
                for a, n, mask, off in recs:
                    node += struct.pack("<II", n, mask) + b"".join(
                        struct.pack("<Q", o) for o in off + [0]) + struct.pack("<Q", a)
                node += struct.pack("<II", 0, 0) + b"".join(
                    struct.pack("<Q", d) for d in list(arr.shape) + [0])
                bt = self._alloc(node)
                layout = struct.pack("<BBBQ", 3, 2, arr.ndim + 1, bt) + b"".join(
                    struct.pack("<I", c) for c in cd + (es,))
            if filters:
                if self.v == 0:
                    fp = struct.pack("<BB6x", 1, len(filters))
                    for fid, cdv in filters:
                        fp += struct.pack("<HHHH", fid, 0, 0, len(cdv)) + b"".join(
                            struct.pack("<I", x) for x in cdv)
                        if len(cdv) % 2:
                            fp += bytes(4)
                else:
                    fp = struct.pack("<BB", 2, len(filters))
                    for fid, cdv in filters:
                        fp += struct.pack("<HHH", fid, 0, len(cdv)) + b"".join(
                            struct.pack("<I", x) for x in cdv)
                msgs.append((0x0B, fp))
        msgs.append((8, layout))
        msgs += [(0x0C, self._attr(k, v)) for k, v in ds.attrs.items()]
        return self._ohdr(msgs)

    def _group(self, g: Group):
        kids = [(name, self._object(obj)) for name, obj in sorted(g.members.items())]
        attrs = [(0x0C, self._attr(k, v)) for k, v in g.attrs.items()]
        if self.v == 2:
            links = [(2, struct.pack("<BBQQ", 0, 0, UNDEF, UNDEF)), (10, bytes([0, 0]))]
            for name, addr in kids:
                nm = name.encode()
                links.append((6, struct.pack("<BBB", 1, 0, len(nm)) + nm + struct.pack("<Q", addr)))
            return self._ohdr(links + attrs), None
        # local heap: offset 0 holds the empty name, then every member name
        heap = bytearray(8)
        offs = []
        for name, _ in kids:
            offs.append(len(heap))
            heap += _pad8(name.encode() + b"\0")
        data = self._alloc(bytes(heap))
        hp = self._alloc(b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(heap), UNDEF, data))
        snods, keys = [], [0]
        per = 8
        for s in range(0, max(len(kids), 1), per):
            part = kids[s:s + per]
            ent = b"".join(struct.pack("<QQII16x", offs[s + j], addr, 0, 0)
                           for j, (_, addr) in enumerate(part))
            snods.append(self._alloc(b"SNOD" + struct.pack("<BBH", 1, 0, len(part)) + ent))
            keys.append(offs[s + len(part) - 1] if part else 0)
        node = b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), UNDEF, UNDEF)
        for i, sn in enumerate(snods):
            node += struct.pack("<QQ", keys[i], sn)
        node += struct.pack("<Q", keys[-1])
        bt = self._alloc(node)
        return self._ohdr([(0x11, struct.pack("<QQ", bt, hp))] + attrs), (bt, hp)

    def _object(self, obj) -> int:
        if isinstance(obj, Group):
            return self._group(obj)[0]
        return self._dataset(obj)

    def write(self, path: str, root: Group) -> None:
        sb_size = 96 if self.v == 0 else 48
        self.buf = bytearray(sb_size)
        addr, scratch = self._group(root)
        eof = len(self.buf)
        base = self.userblock
        if self.v == 0:
            bt, hp = scratch
            sb = (SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", 4, 16, 0) +
                  struct.pack("<QQQQ", base, UNDEF, eof, UNDEF) +
                  struct.pack("<QQII", 0, addr, 1, 0) + struct.pack("<QQ", bt, hp))
        else:
            sb = SIG + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", base, UNDEF, eof, addr)
            sb += struct.pack("<I", lookup3(sb))
        assert len(sb) == sb_size
        self.buf[:sb_size] = sb
        with open(path, "wb") as f:
            f.write(bytes(self.userblock) + bytes(self.buf))


def eman2_stack(images, version: int = 0, **ds_opts) -> Group:
    """EMAN2 layout: /MDF/images/<i> groups with EMAN.nx/ny/nz attributes and
    a dataset "image" of shape (ny, nx) (or (nz, ny, nx))."""
    imgs = {}
    for i, im in enumerate(images):
        im = np.asarray(im)
        nz, ny, nx = (1,) + im.shape if im.ndim == 2 else im.shape
        imgs[str(i)] = Group({"image": Dataset(im, **ds_opts)},
                             {"EMAN.nx": np.int32(nx), "EMAN.ny": np.int32(ny),
                              "EMAN.nz": np.int32(nz), "EMAN.apix_x": np.float32(1.0),
                              "EMAN.micrograph_id": "synthetic"})
    return Group({"MDF": Group({"images": Group(imgs, {"imageid_max": np.int32(len(images) - 1)})})})


def eman1_stack(images, **ds_opts) -> Group:
    """EMAN1 layout: root datasets "0", "1", ... and a root attribute
    num_dataset; the dataspace lists nx first (the reader takes nx = dims[0])."""
    members = {}
    for i, im in enumerate(images):
        im = np.asarray(im)
        members[str(i)] = Dataset(im.reshape(im.shape[::-1]), **ds_opts)
    return Group(members, {"num_dataset": np.int32(len(images))})
