"""Minimal read-only HDF5 parser for the reference's DADF5 result files.

Covers what those files use and nothing more: superblock version 0,
old-style groups (symbol-table message, version-1 B-trees of SNOD nodes,
local heaps), version-1 object headers with continuation blocks, and
datasets with contiguous or chunked layout (layout message version 3,
version-1 chunk B-trees) filtered by deflate, shuffle and Fletcher-32.
Datatypes: little-endian IEEE floats and integers, fixed-length strings.
Attributes are read from object headers.  Used only to extract test
fixtures (arrays of numbers, attribute strings); nothing in the file is
executed.  Written from the published HDF5 file-format specification.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


def fletcher32(b: bytes) -> int:
    """HDF5's Fletcher-32: 16-bit big-endian words (an odd last byte as the
    high byte of a word), both sums folded to 16 bits every 360 words."""
    a = np.frombuffer(b, dtype=np.uint8)
    n = a.size
    words = (a[: n - n % 2].astype(np.uint64).reshape(-1, 2) * np.array([256, 1], np.uint64)).sum(1)
    if n % 2:
        words = np.append(words, np.uint64(int(a[-1]) << 8))
    fold = lambda s: (s & 0xFFFF) + (s >> 16)
    s1 = s2 = 0
    for k in range(0, words.size, 360):
        blk = words[k:k + 360]
        c = np.cumsum(blk, dtype=np.uint64) + np.uint64(s1)
        s1 = fold(int(c[-1]))
        s2 = fold(s2 + int(c.sum(dtype=np.uint64)))
    return (fold(s2) << 16) | fold(s1)


class H5File:
    def __init__(self, path: str):
        with open(path, "rb") as f:
            self.b = f.read()
        if self.b[:8] != SIG:
            raise ValueError("not an HDF5 file")
        ver = self.b[8]
        if ver != 0:
            raise ValueError(f"superblock version {ver} not supported")
        self.so, self.sl = self.b[13], self.b[14]  # sizes of offsets and lengths
        if self.so != 8 or self.sl != 8:
            raise ValueError("only 8-byte offsets and lengths")
        # superblock v0: 8 sig, 4 versions, 1 reserved? -> fields per spec
        p = 24  # after group leaf/internal K (2+2) and consistency flags (4)
        self.base, _free, self.eof, _drv = struct.unpack_from("<4Q", self.b, p)
        p += 32
        # root group symbol table entry
        _name, self.root_oh, cache, _res = struct.unpack_from("<QQII", self.b, p)
        self.root_btree, self.root_heap = struct.unpack_from("<QQ", self.b, p + 24) if cache == 1 else (None, None)

    # -- low level -------------------------------------------------------------
    def u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.b, off)

    def messages(self, oh: int):
        """(type, data bytes) of every message of a version-1 object header."""
        ver, _r, nmsg, _refc, hsize = self.u("BBHII", oh)
        if ver != 1:
            raise ValueError(f"object header version {ver} at {oh}")
        blocks = [(oh + 16, hsize)]  # messages start 8-aligned after the 12-byte prefix
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p, end = start, start + size
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize, _flags = self.u("HHB", p)
                data = self.b[p + 8:p + 8 + msize]
                p += 8 + msize
                if mtype == 0x10:  # continuation
                    coff, clen = struct.unpack_from("<QQ", data, 0)
                    blocks.append((coff, clen))
                out.append((mtype, data))
        return out

    def local_heap_name(self, heap: int, off: int) -> str:
        assert self.b[heap:heap + 4] == b"HEAP"
        data = self.u("Q", heap + 24)[0]
        s = self.b[data + off:self.b.index(b"\x00", data + off)]
        return s.decode()

    def group_entries(self, btree: int, heap: int) -> dict:
        """name -> object header address of an old-style group."""
        out = {}

        def walk(node):
            assert self.b[node:node + 4] == b"TREE", node
            ntype, level, used = self.u("BBH", node + 4)
            assert ntype == 0
            p = node + 24
            p += 8  # key 0
            for _ in range(used):
                child = self.u("Q", p)[0]
                p += 16  # child + next key
                if level > 0:
                    walk(child)
                else:
                    assert self.b[child:child + 4] == b"SNOD"
                    nsym = self.u("H", child + 6)[0]
                    q = child + 8
                    for _ in range(nsym):
                        name_off, oh = self.u("QQ", q)
                        out[self.local_heap_name(heap, name_off)] = oh
                        q += 40

        walk(btree)
        return out

    # -- objects -----------------------------------------------------------------
    def children(self, oh: int) -> dict:
        for t, d in self.messages(oh):
            if t == 0x11:
                bt, hp = struct.unpack_from("<QQ", d, 0)
                return self.group_entries(bt, hp)
        return {}

    def resolve(self, path: str) -> int:
        oh = self.root_oh
        for part in [p for p in path.split("/") if p]:
            ch = self.children(oh)
            if part not in ch:
                raise KeyError(f"{part} not in {sorted(ch)}")
            oh = ch[part]
        return oh

    def keys(self, path: str = "/") -> list:
        return sorted(self.children(self.resolve(path)))

    @staticmethod
    def _dtype(d: bytes):
        cls_ver = d[0]
        cls, size = cls_ver & 0x0F, struct.unpack_from("<I", d, 4)[0]
        bits0 = d[1]
        if cls == 1:  # floating point
            if bits0 & 1:
                raise ValueError("big-endian float")
            return np.dtype(f"<f{size}")
        if cls == 0:  # fixed-point
            if bits0 & 1:
                raise ValueError("big-endian integer")
            signed = bool(bits0 & 0x08)
            return np.dtype(f"<{'i' if signed else 'u'}{size}")
        if cls == 3:  # string
            return np.dtype(f"S{size}")
        raise ValueError(f"datatype class {cls} not supported")

    @staticmethod
    def _dataspace(d: bytes):
        ver, rank, flags = d[0], d[1], d[2]
        p = 8 if ver == 1 else 4
        dims = struct.unpack_from(f"<{rank}Q", d, p)
        return tuple(int(x) for x in dims)

    def attrs(self, path: str) -> dict:
        out = {}
        for t, d in self.messages(self.resolve(path)):
            if t != 0x0C:
                continue
            ver = d[0]
            if ver != 1:
                continue
            nlen, dtlen, dslen = struct.unpack_from("<HHH", d, 2)
            pad = lambda n: (n + 7) & ~7
            p = 8
            name = d[p:p + nlen].split(b"\x00")[0].decode()
            p += pad(nlen)
            try:
                dt = self._dtype(d[p:p + dtlen])
            except ValueError:  # variable-length strings (global heap): not read
                out[name] = None
                continue
            p += pad(dtlen)
            dims = self._dataspace(d[p:p + dslen])
            p += pad(dslen)
            n = int(np.prod(dims)) if dims else 1
            arr = np.frombuffer(d[p:p + n * dt.itemsize], dtype=dt)
            out[name] = arr.reshape(dims) if dims else arr[0]
        return out

    def read(self, path: str) -> np.ndarray:
        msgs = self.messages(self.resolve(path))
        dims = dt = layout = None
        filters = []
        for t, d in msgs:
            if t == 0x01:
                dims = self._dataspace(d)
            elif t == 0x03:
                dt = self._dtype(d)
            elif t == 0x08:
                layout = d
            elif t == 0x0B:
                filters = self._filters(d)
        if layout is None or dt is None:
            raise ValueError(f"{path} is not a dataset")
        ver, cls = layout[0], layout[1]
        if ver != 3:
            raise ValueError(f"layout version {ver}")
        n = int(np.prod(dims)) if dims else 1
        if cls == 1:  # contiguous
            addr, size = struct.unpack_from("<QQ", layout, 2)
            raw = self.b[addr:addr + n * dt.itemsize]
            return np.frombuffer(raw, dtype=dt).reshape(dims).copy()
        if cls == 0:  # compact
            size = struct.unpack_from("<H", layout, 2)[0]
            return np.frombuffer(layout[4:4 + size], dtype=dt).reshape(dims).copy()
        if cls != 2:
            raise ValueError(f"layout class {cls}")
        rank1 = layout[2]
        btree = struct.unpack_from("<Q", layout, 3)[0]
        cdims = struct.unpack_from(f"<{rank1}I", layout, 11)
        chunk = tuple(int(x) for x in cdims[:-1])
        out = np.zeros(dims, dtype=dt)
        for offs, data in self._chunks(btree, rank1 - 1):
            raw = self._unfilter(data, filters, dt.itemsize)
            c = np.frombuffer(raw, dtype=dt)[: int(np.prod(chunk))].reshape(chunk)
            sl = tuple(slice(o, min(o + k, D)) for o, k, D in zip(offs, chunk, dims))
            out[sl] = c[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out

    def _filters(self, d: bytes):
        ver, nf = d[0], d[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(nf):
            fid, nlen, flags, ncv = struct.unpack_from("<HHHH", d, p)
            p += 8
            if ver == 1 or fid >= 256:
                p += (nlen + 7) & ~7 if ver == 1 else nlen
            cv = struct.unpack_from(f"<{ncv}I", d, p)
            p += 4 * ncv
            if ver == 1 and ncv % 2:
                p += 4
            out.append((fid, cv))
        return out

    def _chunks(self, node: int, rank: int):
        assert self.b[node:node + 4] == b"TREE", node
        ntype, level, used = self.u("BBH", node + 4)
        assert ntype == 1
        ksize = 8 + 8 * (rank + 1)
        p = node + 24
        for _ in range(used):
            csize, _mask = self.u("II", p)
            offs = self.u(f"{rank}Q", p + 8)
            child = self.u("Q", p + ksize)[0]
            p += ksize + 8
            if level > 0:
                yield from self._chunks(child, rank)
            else:
                yield tuple(int(o) for o in offs), self.b[child:child + csize]

    @staticmethod
    def _unfilter(data: bytes, filters, itemsize: int) -> bytes:
        for fid, cv in reversed(filters):
            if fid == 3:  # Fletcher-32: checksum in the last 4 bytes, little endian
                if struct.unpack_from("<I", data, len(data) - 4)[0] != fletcher32(data[:-4]):
                    raise ValueError("Fletcher-32 checksum mismatch")
                data = data[:-4]
            elif fid == 1:  # deflate
                data = zlib.decompress(data)
            elif fid == 2:  # shuffle
                size = cv[0] if cv else itemsize
                a = np.frombuffer(data, dtype=np.uint8)
                n = a.size // size
                data = a[: n * size].reshape(size, n).T.copy().tobytes() + bytes(a[n * size:])
            else:
                raise ValueError(f"filter {fid} not supported")
        return data
