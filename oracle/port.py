"""ctypes access to ``oracle/liboracle.so``: this project's C restatement.

Test infrastructure only (tests/, smoke() and bench.py's cpu_baseline leg).
The API mirrors ``oracle.ref`` so a test can run the same checks against the
reference build and against the restatement.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MODE = {"sum": 0, "integrate": 0, "max": 1, "min": 2, "volume": 3}


def available() -> bool:
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_PATH)
        P, S, I = C.c_void_p, C.c_size_t, C.c_int
        L.vto_last_error.restype = C.c_char_p
        L.vto_cloud_create.argtypes = [P, S, P, S, P, P, S, I, I, C.POINTER(P)]
        L.vto_cloud_destroy.argtypes = [P]
        L.vto_cloud_npart.argtypes = [P]
        L.vto_cloud_npart.restype = S
        L.vto_cloud_valid_rgba.argtypes = [P]
        L.vto_cloud_pad.argtypes = [P]
        L.vto_cloud_pad.restype = C.c_double
        L.vto_cloud_orig_ids.argtypes = [P, P]
        L.vto_cloud_num_warnings.argtypes = [P]
        L.vto_cloud_warning.argtypes = [P, I]
        L.vto_cloud_warning.restype = C.c_char_p
        L.vto_tree_nnodes.argtypes = [P]
        L.vto_tree_nnodes.restype = S
        L.vto_tree_export.argtypes = [P] * 8
        L.vto_query.argtypes = [P, P, S, S, P, P]
        L.vto_project.argtypes = [P, P, P, S, I, P, I]
        L.vto_detailed.argtypes = [P, P, P, S, I, I, C.POINTER(P), C.POINTER(S)]
        L.vto_detailed_copy.argtypes = [P] * 6
        L.vto_detailed_destroy.argtypes = [P]
        L.vto_set_mutation.argtypes = [I]
        L.vto_rotate_about.argtypes = [P, S, P, P, P]
        _lib = L
    return _lib


def rotate_about(p, rot, center):
    """(p - center) . rot.T + center in numpy.dot's evaluation order on
    OpenBLAS' x86-64 kernels (vto_rotate_about)."""
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(rot, dtype=np.float64)
    c = np.ascontiguousarray(np.zeros(3) + np.asarray(center), dtype=np.float64)
    out = np.empty_like(p)
    lib().vto_rotate_about(_p(p), len(p), _p(R), _p(c), _p(out))
    return out


def set_mutation(m: int) -> None:
    """Test-only: plant mistake m in the oracle (0 restores it)."""
    lib().vto_set_mutation(int(m))


def _check(rc):
    if rc:
        msg = lib().vto_last_error().decode()
        raise (ValueError if rc == 1 else RuntimeError)(msg)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleCloud:
    """Restated PointCloud (loadPoints + nanoflann-identical buildTree)."""

    def __init__(self, pos, fields, boundbox, vol=None, periodic=False,
                 build_tree=True, nthreads=0):
        L = lib()
        self.nthreads = nthreads
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.ascontiguousarray(fields, dtype=np.float64)
        self.nfields = 1 if f.ndim == 1 else f.shape[1]
        self.fields = f
        sb = np.ascontiguousarray(boundbox, dtype=np.float64)
        v = (np.ascontiguousarray(vol, dtype=np.float64) if vol is not None
             else np.zeros(0))
        h = C.c_void_p()
        _check(L.vto_cloud_create(_p(self.pos), len(self.pos), _p(f),
                                  self.nfields, _p(sb), _p(v), v.size,
                                  int(periodic), int(build_tree), C.byref(h)))
        self.h = h
        self.npart = L.vto_cloud_npart(h)
        self.valid_rgba = bool(L.vto_cloud_valid_rgba(h))
        self.pad = L.vto_cloud_pad(h)
        ids = np.zeros(self.npart, dtype=np.int64)
        L.vto_cloud_orig_ids(h, _p(ids))
        self.orig_ids = ids
        self.warnings = [L.vto_cloud_warning(h, i).decode()
                         for i in range(L.vto_cloud_num_warnings(h))]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.vto_cloud_destroy(self.h)
            self.h = None

    def tree(self):
        """(vind, root_bbox (3,2), nodes) in the same form as
        ``oracle.ref.parse_tree_bytes``."""
        L = lib()
        nn = L.vto_tree_nnodes(self.h)
        vind = np.zeros(self.npart, np.int64)
        feat = np.zeros(nn, np.int32)
        a = np.zeros(nn, np.int64)
        b = np.zeros(nn, np.int64)
        lo = np.zeros(nn)
        hi = np.zeros(nn)
        bb = np.zeros(6)
        L.vto_tree_export(self.h, _p(vind), _p(feat), _p(a), _p(b), _p(lo),
                          _p(hi), _p(bb))
        nodes = [("leaf", int(a[i]), int(b[i])) if feat[i] < 0 else
                 ("inner", int(feat[i]), float(lo[i]), float(hi[i]))
                 for i in range(nn)]
        return vind, bb.reshape(3, 2), nodes

    def query(self, q, k):
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
        ids = np.zeros((len(q), k), dtype=np.int64)
        r2 = np.zeros((len(q), k), dtype=np.float64)
        _check(lib().vto_query(self.h, _p(q), len(q), k, _p(ids), _p(r2)))
        return ids, r2

    def n_out(self, mode):
        return 3 if MODE[mode] == 3 else self.nfields

    def project(self, start, end, mode="sum"):
        s = np.ascontiguousarray(start, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(end, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((len(s), self.n_out(mode)))
        _check(lib().vto_project(self.h, _p(s), _p(e), len(s), MODE[mode],
                                 _p(out), self.nthreads))
        return out

    def detailed(self, start, end, mode="sum"):
        """(values, cell_ids (filtered index), s_enter, s_exit, offsets)."""
        s = np.ascontiguousarray(start, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(end, dtype=np.float64).reshape(-1, 3)
        h, n = C.c_void_p(), C.c_size_t()
        _check(lib().vto_detailed(self.h, _p(s), _p(e), len(s), MODE[mode],
                                  self.nthreads, C.byref(h), C.byref(n)))
        vals = np.zeros((len(s), self.n_out(mode)))
        ids = np.zeros(n.value, dtype=np.int64)
        se = np.zeros(n.value)
        sx = np.zeros(n.value)
        off = np.zeros(len(s) + 1, dtype=np.int64)
        lib().vto_detailed_copy(h, _p(vals), _p(ids), _p(se), _p(sx), _p(off))
        lib().vto_detailed_destroy(h)
        return vals, ids, se, sx, off
