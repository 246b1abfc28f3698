"""Convert the reference's galaxy_interaction snapshot into a numpy fixture.

The reference's golden-image test (tests/test_vortrace.py:12-29 in the
reference) reads ``tests/test_data/galaxy_interaction.hdf5`` through h5py
(tests/conftest.py:8-19): ``PartType0/Coordinates``, ``PartType0/Density`` and
the ``BoxSize`` attribute of ``Parameters``.  h5py is not installed here, so
this script carries a minimal HDF5 reader (superblock v0, v1 object headers,
symbol-table groups, contiguous/chunked float datasets, v1 attributes) that is
just enough for that file.  It writes ``galaxy_interaction_gas.npz`` beside
itself and copies the golden projection ``galaxy_interaction-proj.npy``.

Run in the build container (the reference is mounted there only):
    python tests/golden/make_galaxy_fixture.py
"""
from __future__ import annotations

import hashlib
import os
import struct
import sys
import zlib

import numpy as np

REF_DATA = "/root/reference/tests/test_data"
HERE = os.path.dirname(os.path.abspath(__file__))


class H5:
    """Just enough of the HDF5 file format to read one Arepo snapshot."""

    def __init__(self, path: str):
        with open(path, "rb") as f:
            self.b = f.read()
        if self.b[:8] != b"\x89HDF\r\n\x1a\n":
            raise ValueError("not an HDF5 file")
        if self.b[8] != 0:
            raise ValueError("only superblock version 0 is supported")
        self.so = self.b[13]  # size of offsets
        self.sl = self.b[14]  # size of lengths
        p = 24  # after K values and consistency flags
        p += 4 * self.so  # base, free-space, eof, driver addresses
        self.root = self._symbol_entry(p)

    # -- primitive readers ------------------------------------------------
    def _u(self, p, n):
        return int.from_bytes(self.b[p:p + n], "little")

    def _symbol_entry(self, p):
        name_off = self._u(p, self.so)
        ohdr = self._u(p + self.so, self.so)
        cache = self._u(p + 2 * self.so, 4)
        scratch = p + 2 * self.so + 8
        btree = heap = None
        if cache == 1:
            btree = self._u(scratch, self.so)
            heap = self._u(scratch + self.so, self.so)
        return {"name_off": name_off, "ohdr": ohdr, "btree": btree, "heap": heap}

    def _heap_string(self, heap, off):
        assert self.b[heap:heap + 4] == b"HEAP"
        data = self._u(heap + 8 + 2 * self.sl, self.so)
        s = data + off
        e = self.b.index(b"\x00", s)
        return self.b[s:e].decode()

    def _messages(self, addr):
        """Yield (type, data_offset, size) for every message of a v1 header."""
        if self.b[addr] != 1:
            raise ValueError("only v1 object headers are supported")
        nmsg = self._u(addr + 2, 2)
        hsize = self._u(addr + 8, 4)
        blocks = [(addr + 16, hsize)]
        out = []
        while blocks and len(out) < nmsg:
            p, size = blocks.pop(0)
            end = p + size
            while p + 8 <= end and len(out) < nmsg:
                mtype = self._u(p, 2)
                msize = self._u(p + 2, 2)
                out.append((mtype, p + 8, msize))
                if mtype == 0x10:  # continuation
                    blocks.append((self._u(p + 8, self.so),
                                   self._u(p + 8 + self.so, self.sl)))
                p += 8 + msize
        return out

    # -- groups -------------------------------------------------------------
    def _group_links(self, btree, heap):
        links = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            assert self.b[node:node + 4] == b"TREE"
            level = self.b[node + 5]
            used = self._u(node + 6, 2)
            p = node + 8 + 2 * self.so  # after siblings
            p += self.sl  # key 0
            for _ in range(used):
                child = self._u(p, self.so)
                p += self.so + self.sl
                if level > 0:
                    stack.append(child)
                    continue
                assert self.b[child:child + 4] == b"SNOD"
                nsym = self._u(child + 6, 2)
                q = child + 8
                for _ in range(nsym):
                    ent = self._symbol_entry(q)
                    links[self._heap_string(heap, ent["name_off"])] = ent
                    q += 2 * self.so + 24
        return links

    def _lookup(self, path):
        ent = self.root
        for part in [x for x in path.split("/") if x]:
            btree, heap = ent["btree"], ent["heap"]
            if btree is None:
                for mtype, p, _ in self._messages(ent["ohdr"]):
                    if mtype == 0x11:
                        btree = self._u(p, self.so)
                        heap = self._u(p + self.so, self.so)
            ent = self._group_links(btree, heap)[part]
        return ent

    # -- datatypes / dataspaces ---------------------------------------------
    def _dtype(self, p):
        cls = self.b[p] & 0x0F
        bits0 = self.b[p + 1]
        size = self._u(p + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        if cls == 0:
            return np.dtype(f"{order}{'i' if bits0 & 8 else 'u'}{size}")
        raise ValueError(f"unsupported datatype class {cls}")

    def _dspace(self, p):
        ver, ndim, flags = self.b[p], self.b[p + 1], self.b[p + 2]
        q = p + (8 if ver == 1 else 4)
        return tuple(self._u(q + i * self.sl, self.sl) for i in range(ndim))

    # -- datasets and attributes ----------------------------------------------
    def dataset(self, path):
        ent = self._lookup(path)
        dtype = shape = layout = None
        filters = []
        for mtype, p, _ in self._messages(ent["ohdr"]):
            if mtype == 0x01:
                shape = self._dspace(p)
            elif mtype == 0x03:
                dtype = self._dtype(p)
            elif mtype == 0x08:
                layout = p
            elif mtype == 0x0B:
                filters = self._filters(p)
        ver, cls = self.b[layout], self.b[layout + 1]
        if ver != 3:
            raise ValueError("only v3 data layout messages are supported")
        n = int(np.prod(shape))
        if cls == 1:
            addr = self._u(layout + 2, self.so)
            raw = self.b[addr:addr + n * dtype.itemsize]
            return np.frombuffer(raw, dtype=dtype).reshape(shape).astype(
                dtype.newbyteorder("="))
        if cls == 2:
            nd = self.b[layout + 2]
            btree = self._u(layout + 3, self.so)
            cdims = [self._u(layout + 3 + self.so + 4 * i, 4) for i in range(nd)]
            return self._chunked(btree, shape, cdims[:-1], dtype, filters)
        raise ValueError(f"unsupported layout class {cls}")

    def _filters(self, p):
        ver, nf = self.b[p], self.b[p + 1]
        q = p + 8
        out = []
        for _ in range(nf):
            fid = self._u(q, 2)
            nlen = self._u(q + 2, 2)
            ncv = self._u(q + 6, 2)
            q += 8 + ((nlen + 7) // 8) * 8 + 4 * ncv
            if ncv % 2:
                q += 4
            out.append(fid)
        return out

    def _chunked(self, btree, shape, cdims, dtype, filters):
        out = np.zeros(shape, dtype=dtype.newbyteorder("="))
        nd = len(shape)
        stack = [btree]
        while stack:
            node = stack.pop()
            assert self.b[node:node + 4] == b"TREE"
            level = self.b[node + 5]
            used = self._u(node + 6, 2)
            p = node + 8 + 2 * self.so
            ksize = 8 + 8 * (nd + 1)
            for _ in range(used):
                csize = self._u(p, 4)
                offs = [self._u(p + 8 + 8 * i, 8) for i in range(nd)]
                child = self._u(p + ksize, self.so)
                p += ksize + self.so
                if level > 0:
                    stack.append(child)
                    continue
                raw = self.b[child:child + csize]
                for fid in reversed(filters):
                    if fid == 1:
                        raw = zlib.decompress(raw)
                    elif fid == 2:
                        a = np.frombuffer(raw, np.uint8)
                        raw = a.reshape(dtype.itemsize, -1).T.tobytes()
                    else:
                        raise ValueError(f"unsupported filter {fid}")
                chunk = np.frombuffer(raw, dtype=dtype).reshape(cdims)
                sl = tuple(slice(o, min(o + c, s))
                           for o, c, s in zip(offs, cdims, shape))
                out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out

    def attribute(self, path, name):
        ent = self._lookup(path)
        for mtype, p, _ in self._messages(ent["ohdr"]):
            if mtype != 0x0C:
                continue
            nlen = self._u(p + 2, 2)
            tlen = self._u(p + 4, 2)
            slen = self._u(p + 6, 2)
            q = p + 8
            aname = self.b[q:q + nlen].split(b"\x00")[0].decode()
            q += ((nlen + 7) // 8) * 8
            if aname != name:
                continue
            dtype = self._dtype(q)
            q += ((tlen + 7) // 8) * 8
            shape = self._dspace(q)
            q += ((slen + 7) // 8) * 8
            n = int(np.prod(shape)) if shape else 1
            return np.frombuffer(self.b[q:q + n * dtype.itemsize],
                                 dtype=dtype).reshape(shape or ())
        raise KeyError(name)


def main():
    f = H5(os.path.join(REF_DATA, "galaxy_interaction.hdf5"))
    pos = np.ascontiguousarray(f.dataset("PartType0/Coordinates"))
    dens = np.ascontiguousarray(f.dataset("PartType0/Density"))
    box = float(f.attribute("Parameters", "BoxSize"))
    print("Coordinates", pos.shape, pos.dtype, "Density", dens.shape,
          dens.dtype, "BoxSize", box)
    out = os.path.join(HERE, "galaxy_interaction_gas.npz")
    np.savez_compressed(out, pos=pos, dens=dens, box_size=np.float64(box))
    proj = np.load(os.path.join(REF_DATA, "galaxy_interaction-proj.npy"))
    np.save(os.path.join(HERE, "galaxy_interaction-proj.npy"), proj)
    for name in ("galaxy_interaction_gas.npz", "galaxy_interaction-proj.npy"):
        with open(os.path.join(HERE, name), "rb") as g:
            print(name, hashlib.sha256(g.read()).hexdigest())
    return 0


if __name__ == "__main__":
    sys.exit(main())
