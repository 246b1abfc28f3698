"""ctypes access to ``oracle/_ref/libvtref.so``: the reference's own C++ core.

Test infrastructure only.  The library is the reference's src/*.cpp compiled
unchanged (see oracle/Makefile) behind ``ref_shim.cpp``; it exists where the
build step could see the reference checkout and travels with the tree as a
prebuilt file.  Use ``available()`` before relying on it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libvtref.so")
_lib = None

MODE = {"sum": 0, "integrate": 0, "max": 1, "min": 2, "volume": 3}


def available() -> bool:
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_PATH)
        P, S, I = C.c_void_p, C.c_size_t, C.c_int
        L.vtref_last_error.restype = C.c_char_p
        L.vtref_cloud_create.argtypes = [P, S, P, S, P, P, S, I, I, C.POINTER(P)]
        L.vtref_cloud_destroy.argtypes = [P]
        L.vtref_cloud_npart.argtypes = [P]
        L.vtref_cloud_npart.restype = S
        L.vtref_cloud_valid_rgba.argtypes = [P]
        L.vtref_cloud_pad.argtypes = [P]
        L.vtref_cloud_pad.restype = C.c_double
        L.vtref_cloud_orig_ids.argtypes = [P, P]
        L.vtref_query.argtypes = [P, P, S, S, P, P]
        L.vtref_tree_bytes.argtypes = [P, C.POINTER(P), C.POINTER(S)]
        L.vtref_free.argtypes = [P]
        L.vtref_project.argtypes = [P, P, P, S, I, P]
        L.vtref_detailed.argtypes = [P, P, P, S, I, C.POINTER(P), C.POINTER(S)]
        L.vtref_detailed_copy.argtypes = [P, P, P, P, P, P]
        L.vtref_detailed_destroy.argtypes = [P]
        L.vtref_num_warnings.restype = S
        L.vtref_warning.restype = C.c_char_p
        L.vtref_warning.argtypes = [S]
        L.vtref_capture_warnings()
        _lib = L
    return _lib


def _check(rc):
    if rc:
        msg = lib().vtref_last_error().decode()
        raise (ValueError if rc == 1 else RuntimeError)(msg)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class RefCloud:
    """The reference's PointCloud (loadPoints + buildTree)."""

    def __init__(self, pos, fields, boundbox, vol=None, periodic=False,
                 build_tree=True):
        L = lib()
        L.vtref_clear_warnings()
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.ascontiguousarray(fields, dtype=np.float64)
        self.nfields = 1 if f.ndim == 1 else f.shape[1]
        self.fields = f
        sb = np.ascontiguousarray(boundbox, dtype=np.float64)
        v = (np.ascontiguousarray(vol, dtype=np.float64) if vol is not None
             else np.zeros(0))
        h = C.c_void_p()
        _check(L.vtref_cloud_create(_p(self.pos), len(self.pos), _p(f),
                                    self.nfields, _p(sb), _p(v), v.size,
                                    int(periodic), int(build_tree), C.byref(h)))
        self.h = h
        self.npart = L.vtref_cloud_npart(h)
        self.valid_rgba = bool(L.vtref_cloud_valid_rgba(h))
        self.pad = L.vtref_cloud_pad(h)
        ids = np.zeros(self.npart, dtype=np.int64)
        L.vtref_cloud_orig_ids(h, _p(ids))
        self.orig_ids = ids
        self.warnings = [L.vtref_warning(i).decode()
                         for i in range(L.vtref_num_warnings())]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.vtref_cloud_destroy(self.h)
            self.h = None

    def query(self, q, k):
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
        ids = np.zeros((len(q), k), dtype=np.int64)
        r2 = np.zeros((len(q), k), dtype=np.float64)
        _check(lib().vtref_query(self.h, _p(q), len(q), k, _p(ids), _p(r2)))
        return ids, r2

    def tree_bytes(self) -> bytes:
        buf, n = C.c_void_p(), C.c_size_t()
        _check(lib().vtref_tree_bytes(self.h, C.byref(buf), C.byref(n)))
        out = C.string_at(buf, n.value)
        lib().vtref_free(buf)
        return out

    def n_out(self, mode):
        return 3 if MODE[mode] == 3 else self.nfields

    def project(self, start, end, mode="sum"):
        s = np.ascontiguousarray(start, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(end, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((len(s), self.n_out(mode)))
        _check(lib().vtref_project(self.h, _p(s), _p(e), len(s), MODE[mode],
                                   _p(out)))
        return out

    def detailed(self, start, end, mode="sum"):
        """(values, cell_ids (filtered index), s_enter, s_exit, offsets)."""
        s = np.ascontiguousarray(start, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(end, dtype=np.float64).reshape(-1, 3)
        h, n = C.c_void_p(), C.c_size_t()
        _check(lib().vtref_detailed(self.h, _p(s), _p(e), len(s), MODE[mode],
                                    C.byref(h), C.byref(n)))
        vals = np.zeros((len(s), self.n_out(mode)))
        ids = np.zeros(n.value, dtype=np.int64)
        se = np.zeros(n.value)
        sx = np.zeros(n.value)
        off = np.zeros(len(s) + 1, dtype=np.int64)
        lib().vtref_detailed_copy(h, _p(vals), _p(ids), _p(se), _p(sx), _p(off))
        lib().vtref_detailed_destroy(h)
        return vals, ids, se, sx, off


def parse_tree_bytes(b: bytes):
    """Decode a nanoflann v1.3.2 index dump (include/nanoflann.hpp:1023-1071).

    Layout: m_size (size_t), dim (int), root_bbox (3 x {low, high} f64),
    m_leaf_max_size (size_t), vind (size_t count + entries), then the nodes in
    pre-order, 40 bytes each: a 24-byte union ({left, right} size_t for a leaf,
    {divfeat int, pad, divlow f64, divhigh f64} for an inner node) followed by
    the child1/child2 pointers, which act only as null / non-null flags.
    Returns (vind, root_bbox (3,2), nodes) with nodes a list of tuples
    ('leaf', left, right) or ('inner', divfeat, divlow, divhigh) in pre-order.
    """
    p = 0

    def take(fmt):
        nonlocal p
        v = np.frombuffer(b, dtype=fmt, count=1, offset=p)[0]
        p += np.dtype(fmt).itemsize
        return v

    m_size = int(take("<u8"))
    take("<i4")  # dim, written as a bare 4-byte int
    bbox = np.frombuffer(b, dtype="<f8", count=6, offset=p).reshape(3, 2)
    p += 48
    take("<u8")  # leaf size
    nv = int(take("<u8"))
    vind = np.frombuffer(b, dtype="<u8", count=nv, offset=p).astype(np.int64)
    p += 8 * nv
    assert nv == m_size
    nodes = []
    stack = [1]
    while stack:
        stack.pop()
        raw = b[p:p + 40]
        p += 40
        c1 = int.from_bytes(raw[24:32], "little")
        c2 = int.from_bytes(raw[32:40], "little")
        if c1 == 0 and c2 == 0:
            left = int.from_bytes(raw[0:8], "little")
            right = int.from_bytes(raw[8:16], "little")
            nodes.append(("leaf", left, right))
        else:
            feat = int.from_bytes(raw[0:4], "little", signed=True)
            lo = np.frombuffer(raw, "<f8", 1, 8)[0]
            hi = np.frombuffer(raw, "<f8", 1, 16)[0]
            nodes.append(("inner", feat, float(lo), float(hi)))
            stack.append(2)
            stack.append(1)
    assert p == len(b), (p, len(b))
    return vind, bbox.copy(), nodes
