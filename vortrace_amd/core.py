"""Low-level classes mirroring the reference's ``Cvortrace`` module.

Same class, method and argument names as src/bindings.cpp:71-326
(``PointCloud``, ``Projection``, ``ReductionMode``, ``set_verbose``,
``get_verbose``) over the C-ABI of include/vortrace_hip.h.  Inputs are numpy
arrays or torch tensors; ray and query arrays go to the device the cloud was
built on; ``return*`` methods give numpy arrays like the reference.  The
``*_tensor`` methods keep everything on the device.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np
import torch

from . import _lib

_P = C.c_void_p


class ReductionMode(enum.IntEnum):
    """include/mytypes.hpp:14 (values are part of the ABI)."""
    Sum = 0
    Max = 1
    Min = 2
    VolumeRender = 3


Sum, Max, Min, VolumeRender = (ReductionMode.Sum, ReductionMode.Max,
                               ReductionMode.Min, ReductionMode.VolumeRender)


def set_verbose(verbose: bool) -> None:
    _lib.lib().vt_set_verbose(int(bool(verbose)))


def get_verbose() -> bool:
    return bool(_lib.lib().vt_get_verbose())


def _host_f64(a, name):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _stream():
    return _P(torch.cuda.current_stream().cuda_stream)


# Bytes of one segment record of vt_project_recorded (VT_SEGREC_BYTES).
SEGREC_BYTES = 32


def _check_counts(counts: torch.Tensor, ngrid: int) -> None:
    if counts.dtype != torch.int64 or counts.numel() != ngrid or not counts.is_contiguous():
        raise ValueError("counts must be a contiguous int64 tensor of ngrid entries")


def device_f64(a, device) -> torch.Tensor:
    """A contiguous FP64 tensor on `device` (copied only when needed)."""
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=device, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)


class PointCloud:
    """Particle positions, fields and the device kd-tree (src/pointcloud.cpp)."""

    def __init__(self):
        self._h = None
        self._device = None
        _lib.lib()

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and _lib._lib is not None:
            _lib._lib.vt_cloud_destroy(h)

    def _handle(self):
        if self._h is None:
            raise RuntimeError("PointCloud has no points loaded")
        return self._h

    # -- construction --------------------------------------------------------
    def loadPoints(self, pos, fields_in, subbox, vol=None, periodic=False):
        """PointCloud.loadPoints (src/bindings.cpp:29-69,112-116)."""
        pos = _host_f64(pos, "pos")
        fields = _host_f64(fields_in, "fields_in")
        if pos.ndim != 2:
            raise ValueError("pos array must be two-dimensional")
        if fields.ndim == 1:
            nf, nfp = 1, fields.shape[0]
        elif fields.ndim == 2:
            nf, nfp = fields.shape[1], fields.shape[0]
        else:
            raise ValueError("fields array must be one-dimensional or two-dimensional")
        if pos.shape[0] != nfp:
            raise ValueError(f"Input sizes must match: pos has {pos.shape[0]} "
                             f"rows but fields has {nfp}")
        if pos.shape[1] != 3:
            raise ValueError("pos must have shape (N, 3)")
        sb = np.ascontiguousarray(np.asarray(subbox, dtype=np.float64).ravel())
        if sb.size != 6:
            raise ValueError("subbox must have 6 entries")
        v = (np.ascontiguousarray(_host_f64(vol, "vol").ravel())
             if vol is not None else np.zeros(0))
        h = _P()
        L = _lib.lib()
        self._fields_dev = None
        _lib.check(L.vt_cloud_create(_ptr(pos), pos.shape[0], _ptr(fields), nf,
                                     _ptr(sb), _ptr(v) if v.size else None,
                                     v.size, int(bool(periodic)), C.byref(h)))
        old, self._h = self._h, h
        if old is not None:
            L.vt_cloud_destroy(old)
        self._pos = pos
        self._fields = fields.reshape(pos.shape[0], nf)

    def buildTree(self):
        """PointCloud.buildTree: host build, then upload to the current device."""
        torch.cuda.current_device()  # initialise torch's HIP context first
        _lib.check(_lib.lib().vt_cloud_build_tree(self._handle()))
        self._device = torch.device("cuda", torch.cuda.current_device())

    def build_host_tree(self):
        """Host half of buildTree only (no device needed)."""
        _lib.check(_lib.lib().vt_cloud_build_host_tree(self._handle()))

    # -- getters (src/bindings.cpp:134-154) -------------------------------------
    def get_nfields(self) -> int:
        return int(_lib.lib().vt_cloud_nfields(self._handle()))

    def get_npart(self) -> int:
        return int(_lib.lib().vt_cloud_npart(self._handle()))

    def get_periodic(self) -> bool:
        return bool(_lib.lib().vt_cloud_periodic(self._handle()))

    def get_valid_rgba(self) -> bool:
        return bool(_lib.lib().vt_cloud_valid_rgba(self._handle()))

    def get_tree_built(self) -> bool:
        return bool(_lib.lib().vt_cloud_tree_built(self._handle()))

    def get_pad(self) -> float:
        return float(_lib.lib().vt_cloud_pad(self._handle()))

    def get_subbox(self):
        out = np.zeros(6)
        _lib.lib().vt_cloud_subbox(self._handle(), _ptr(out))
        return [float(x) for x in out]

    def get_orig_ids(self) -> np.ndarray:
        out = np.zeros(self.get_npart(), dtype=np.int64)
        _lib.lib().vt_cloud_orig_ids(self._handle(), _ptr(out))
        return out

    def get_pt(self, idx: int):
        return [float(x) for x in self._pos[self.get_orig_ids()[idx]]]

    def get_field(self, idx: int, f: int) -> float:
        fd = getattr(self, "_fields_dev", None)
        if fd is not None:
            return float(fd[idx, f].item())
        return float(self._fields[self.get_orig_ids()[idx], f])

    def tree_info(self) -> dict:
        ti = _lib.TreeInfo()
        _lib.check(_lib.lib().vt_cloud_tree_info(self._handle(), C.byref(ti)))
        return {"nodes": ti.nodes, "leaves": ti.leaves,
                "max_depth": ti.max_depth, "build_seconds": ti.build_seconds}

    def tree_export(self):
        """(vind, root_bbox (3,2), nodes) in pre-order; nodes are
        ('leaf', left, right) or ('inner', divfeat, divlow, divhigh)."""
        ti = self.tree_info()
        n, nn = self.get_npart(), ti["nodes"]
        vind = np.zeros(n, np.int64)
        feat = np.zeros(nn, np.int32)
        a = np.zeros(nn, np.int64)
        b = np.zeros(nn, np.int64)
        lo = np.zeros(nn)
        hi = np.zeros(nn)
        bb = np.zeros(6)
        _lib.check(_lib.lib().vt_cloud_tree_export(
            self._handle(), _ptr(vind), _ptr(feat), _ptr(a), _ptr(b), _ptr(lo),
            _ptr(hi), _ptr(bb)))
        nodes = [("leaf", int(a[i]), int(b[i])) if feat[i] < 0 else
                 ("inner", int(feat[i]), float(lo[i]), float(hi[i]))
                 for i in range(nn)]
        return vind, bb.reshape(3, 2), nodes

    # -- tree bytes (src/bindings.cpp:125-132) ------------------------------------
    def saveTreeToBytes(self) -> bytes:
        """The tree as nanoflann 1.3.2's index dump (vt_cloud_save_tree)."""
        L = _lib.lib()
        size = C.c_size_t()
        _lib.check(L.vt_cloud_save_tree(self._handle(), None, 0, C.byref(size)))
        buf = (C.c_uint8 * max(1, size.value))()
        _lib.check(L.vt_cloud_save_tree(self._handle(), buf, size.value, C.byref(size)))
        return bytes(buf)[: size.value]

    def loadTreeFromBytes(self, data: bytes) -> None:
        """Replaces the tree with a nanoflann index dump and uploads it to the
        current device (vt_cloud_load_tree)."""
        data = bytes(data)
        buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
        torch.cuda.current_device()
        _lib.check(_lib.lib().vt_cloud_load_tree(self._handle(), buf, len(data), 1))
        self._device = torch.device("cuda", torch.cuda.current_device())

    def saveTree(self, filename: str) -> None:
        """PointCloud::saveTree (src/pointcloud.cpp:214-225): the tree as
        nanoflann's index dump in a file."""
        data = self.saveTreeToBytes()
        try:
            f = open(filename, "wb")
        except OSError:
            raise RuntimeError(f"Cannot open file for writing: {filename}") from None
        with f:
            f.write(data)

    def loadTree(self, filename: str) -> None:
        """PointCloud::loadTree (src/pointcloud.cpp:227-240): replaces the
        tree with the index dump in a file and uploads it to the current
        device."""
        if self._h is None or self.get_npart() == 0:
            raise RuntimeError("Cannot load tree: no points loaded")
        try:
            f = open(filename, "rb")
        except OSError:
            raise RuntimeError(f"Cannot open file for reading: {filename}") from None
        with f:
            data = f.read()
        self.loadTreeFromBytes(data)

    def load_host_tree_bytes(self, data: bytes) -> None:
        """loadTreeFromBytes without the upload (no device needed)."""
        data = bytes(data)
        buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
        _lib.check(_lib.lib().vt_cloud_load_tree(self._handle(), buf, len(data), 0))

    # -- fields on the device ----------------------------------------------------
    def set_ray_rows(self, row: int) -> None:
        """Declares the coming calls' rays to be the pixels of row-major
        images with `row` rays per row (0: no image order), so the kernels
        take them in 8x8-pixel tiles (vt_cloud_set_ray_rows).  Speed only."""
        row = int(row)
        _lib.check(_lib.lib().vt_cloud_set_ray_rows(self._handle(), row))
        self._ray_rows = row

    def get_ray_rows(self) -> int:
        return getattr(self, "_ray_rows", 0)

    def set_ray_order(self, order: torch.Tensor | None) -> None:
        """Order in which the kernels take the rays of coming calls with
        len(order) rays: a permutation of [0, n) as an int32 device tensor,
        or None to clear (vt_cloud_set_ray_order).  Speed only."""
        if order is None:
            _lib.check(_lib.lib().vt_cloud_set_ray_order(self._handle(), None, 0, _stream()))
            self._ray_order_n = 0
            return
        if not isinstance(order, torch.Tensor) or order.dtype != torch.int32 \
                or order.dim() != 1 or order.device != self._device:
            raise ValueError("order must be a 1-D int32 tensor on the cloud's device")
        order = order.contiguous()
        self._ray_order_n = 0  # a rejected order clears the previous one
        _lib.check(_lib.lib().vt_cloud_set_ray_order(self._handle(), _P(order.data_ptr()),
                                                     order.numel(), _stream()))
        self._ray_order_n = order.numel()

    def order_rays(self, start: torch.Tensor | None = None, end: torch.Tensor | None = None,
                   counts: torch.Tensor | None = None, row: int = 0, block: int = 1 << 15,
                   out: bool = False, keys: torch.Tensor | None = None):
        """Sets the order in which the kernels take the coming calls' rays
        (as set_ray_order) to longest first, built on the device
        (vt_cloud_order_rays): each block of `block` fetch slots of a ray
        partition takes its runs of 64 slots by descending key, the largest
        of `counts` (int64 segment counts of a previous pass over the same
        rays) over the run's rays, or, without counts, a cost estimated from
        the rays (start, end) and the point density.  `row`: rays per image
        row of row-major images (8x8-pixel tiles), 0 for index order.
        Returns the order (int32 device tensor) when `out`; `keys` (int64,
        one per run) receives the run keys.  Speed only."""
        if not self.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        dev = self._device
        if counts is not None:
            if counts.dtype != torch.int64 or counts.dim() != 1 or counts.device != dev:
                raise ValueError("counts must be a 1-D int64 tensor on the cloud's device")
            counts = counts.contiguous()
            n = counts.numel()
            sp = ep = None
        else:
            start = device_f64(start, dev).reshape(-1, 3)
            end = device_f64(end, dev).reshape(-1, 3)
            if start.shape != end.shape:
                raise ValueError("start and end must have the same shape")
            n = start.shape[0]
            sp, ep = _P(start.data_ptr()), _P(end.data_ptr())
        order = torch.empty(n, dtype=torch.int32, device=dev) if out else None
        self._ray_order_n = 0
        _lib.check(_lib.lib().vt_cloud_order_rays(
            self._handle(), sp, ep, _P(counts.data_ptr()) if counts is not None else None,
            n, int(row), int(block), _P(order.data_ptr()) if out else None,
            _P(keys.data_ptr()) if keys is not None else None, _stream()))
        self._ray_order_n = n
        return order

    def get_ray_order_size(self) -> int:
        """Ray count of the order set by set_ray_order (0: none)."""
        return getattr(self, "_ray_order_n", 0)

    def set_fields_tensor(self, fields_filtered: torch.Tensor) -> None:
        """Replaces the fields with a device tensor in filtered point order
        (n, nfields), keeping the tree (vt_cloud_set_fields)."""
        if not self.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        nf = self.get_nfields()
        f = device_f64(fields_filtered, self._device).reshape(-1)
        if f.numel() != self.get_npart() * nf:
            raise ValueError(f"fields must have {self.get_npart()} x {nf} values, "
                             f"got {f.numel()}")
        _lib.check(_lib.lib().vt_cloud_set_fields(self._handle(), _P(f.data_ptr()),
                                                  _stream()))
        self._fields_dev = f.reshape(self.get_npart(), nf)

    def lanes(self, reduction=Sum, recorded: bool = False) -> int:
        """Resident lanes of the forward kernel a projection with `reduction`
        (recording its segments when `recorded`) launches on this cloud
        (vt_cloud_lanes)."""
        n = C.c_int64()
        _lib.check(_lib.lib().vt_cloud_lanes(self._handle(), int(reduction),
                                             int(bool(recorded)), C.byref(n)))
        return int(n.value)

    # -- queries -----------------------------------------------------------------
    def query_tensor(self, q: torch.Tensor, k: int):
        """k-NN of device-resident queries (nq, 3) -> (ids int64, r2 f64)."""
        if not self.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        q = device_f64(q, self._device).reshape(-1, 3)
        ids = torch.empty((q.shape[0], k), dtype=torch.int64, device=self._device)
        r2 = torch.empty((q.shape[0], k), dtype=torch.float64, device=self._device)
        _lib.check(_lib.lib().vt_query_knn(self._handle(), _P(q.data_ptr()),
                                           q.shape[0], int(k), _P(ids.data_ptr()),
                                           _P(r2.data_ptr()), _stream()))
        return ids, r2

    def queryTree(self, query_pt, k=None):
        """PointCloud.queryTree (src/bindings.cpp:155-174): the nearest id, or
        (ids uint64[k], r2 f64[k]) when k is given."""
        q = np.asarray(query_pt, dtype=np.float64).reshape(1, 3)
        ids, r2 = self.query_tensor(q, 1 if k is None else int(k))
        ids = ids.cpu().numpy()[0]
        if k is None:
            return int(ids[0])
        return ids.astype(np.uint64), r2.cpu().numpy()[0]


class Projection:
    """Rays for one projection (src/projection.cpp, binding 176-238)."""

    def __init__(self, pos_start, pos_end):
        if not isinstance(pos_start, torch.Tensor):
            pos_start = np.asarray(pos_start, dtype=np.float64)
        if not isinstance(pos_end, torch.Tensor):
            pos_end = np.asarray(pos_end, dtype=np.float64)
        if pos_start.ndim != 2 or pos_end.ndim != 2:
            raise ValueError("PROJECTION: pos_start and pos_end array must both be "
                             "two-dimensional\n")
        ns = int(np.prod(tuple(pos_start.shape)))
        ne = int(np.prod(tuple(pos_end.shape)))
        if ns != ne:
            raise ValueError("PROJECTION: pos_start and pos_end must have same sizes")
        self._start = pos_start
        self._end = pos_end
        self.ngrid = ns // 3
        self._data = None
        self._segments = None

    def _rays(self, cloud: PointCloud):
        dev = cloud._device
        s = device_f64(self._start, dev).reshape(-1)[: 3 * self.ngrid].reshape(-1, 3)
        e = device_f64(self._end, dev).reshape(-1)[: 3 * self.ngrid].reshape(-1, 3)
        return s.contiguous(), e.contiguous()

    @staticmethod
    def _nout(cloud, mode):
        return 3 if int(mode) == VolumeRender else cloud.get_nfields()

    def project_tensor(self, cloud: PointCloud, reduction=Sum,
                       counts: torch.Tensor | None = None) -> torch.Tensor:
        """makeProjection with the result left on the device (ngrid, n_out);
        also writes each ray's segment count into `counts` (int64, ngrid)
        when given."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        s, e = self._rays(cloud)
        out = torch.zeros((self.ngrid, self._nout(cloud, reduction)),
                          dtype=torch.float64, device=cloud._device)
        cp = None
        if counts is not None:
            _check_counts(counts, self.ngrid)
            cp = _P(counts.data_ptr())
        _lib.check(_lib.lib().vt_project_counted(cloud._handle(), _P(s.data_ptr()),
                                                 _P(e.data_ptr()), self.ngrid,
                                                 int(reduction), _P(out.data_ptr()), cp,
                                                 _stream()))
        return out

    def project_recorded_tensor(self, cloud: PointCloud, reduction, counts: torch.Tensor,
                                records: torch.Tensor):
        """project_tensor for 'sum' or 'volume' that also records every
        segment into `records` (uint8, (capacity, SEGREC_BYTES), on the
        device) for adjoint_recorded_tensor (vt_project_recorded).  Returns
        (image, records used), records used -1 when the pool was too small."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        _check_counts(counts, self.ngrid)
        if records.dtype != torch.uint8 or records.dim() != 2 or \
                records.shape[1] != SEGREC_BYTES or not records.is_contiguous():
            raise ValueError(f"records must be a contiguous uint8 tensor (capacity, "
                             f"{SEGREC_BYTES})")
        s, e = self._rays(cloud)
        out = torch.zeros((self.ngrid, self._nout(cloud, reduction)),
                          dtype=torch.float64, device=cloud._device)
        used = C.c_size_t(0)
        _lib.check(_lib.lib().vt_project_recorded(
            cloud._handle(), _P(s.data_ptr()), _P(e.data_ptr()), self.ngrid, int(reduction),
            _P(out.data_ptr()), _P(counts.data_ptr()), _P(records.data_ptr()),
            records.shape[0], C.byref(used), _stream()))
        n = used.value
        return out, (-1 if n == C.c_size_t(-1).value else n)

    def adjoint_recorded_tensor(self, cloud: PointCloud, grad_out: torch.Tensor, reduction,
                                counts: torch.Tensor, records: torch.Tensor,
                                n_records: int) -> torch.Tensor:
        """adjoint_tensor from the records of project_recorded_tensor
        (vt_adjoint_recorded); the same bits as adjoint_tensor."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        _check_counts(counts, self.ngrid)
        if n_records < 0:
            raise ValueError("the recorded projection overflowed its record pool")
        if n_records > records.shape[0]:
            raise ValueError("n_records exceeds the pool")
        g = device_f64(grad_out, cloud._device).reshape(-1)
        nout = self._nout(cloud, reduction)
        if g.numel() != self.ngrid * nout:
            raise ValueError(f"grad_out must have {self.ngrid} x {nout} values, got {g.numel()}")
        out = torch.empty((cloud.get_npart(), cloud.get_nfields()), dtype=torch.float64,
                          device=cloud._device)
        _lib.check(_lib.lib().vt_adjoint_recorded(
            cloud._handle(), self.ngrid, int(reduction), _P(counts.data_ptr()),
            _P(records.data_ptr()), n_records, _P(g.data_ptr()), _P(out.data_ptr()),
            _stream()))
        return out

    def makeProjection(self, cloud: PointCloud, reduction=Sum):
        self._data = self.project_tensor(cloud, reduction)

    def returnProjection(self) -> np.ndarray:
        if self._data is None:
            raise RuntimeError("Projection has not yet been made")
        d = self._data.cpu().numpy()
        return d[:, 0].copy() if d.shape[1] == 1 else d

    def segments_tensor(self, cloud: PointCloud, reduction=Sum):
        """makeDetailedProjection on the device: (values, cell_ids, s_enter,
        s_exit, offsets) as device tensors (ids are filtered indices)."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        s, e = self._rays(cloud)
        dev = cloud._device
        L = _lib.lib()
        counts = torch.zeros(self.ngrid, dtype=torch.int64, device=dev)
        _lib.check(L.vt_project_count(cloud._handle(), _P(s.data_ptr()),
                                      _P(e.data_ptr()), self.ngrid,
                                      _P(counts.data_ptr()), _stream()))
        offsets = torch.zeros(self.ngrid + 1, dtype=torch.int64, device=dev)
        torch.cumsum(counts, 0, out=offsets[1:])
        total = int(offsets[-1].item())
        vals = torch.zeros((self.ngrid, self._nout(cloud, reduction)),
                           dtype=torch.float64, device=dev)
        ids = torch.empty(max(total, 1), dtype=torch.int64, device=dev)
        se = torch.empty(max(total, 1), dtype=torch.float64, device=dev)
        sx = torch.empty(max(total, 1), dtype=torch.float64, device=dev)
        _lib.check(L.vt_project_segments(
            cloud._handle(), _P(s.data_ptr()), _P(e.data_ptr()), self.ngrid,
            int(reduction), _P(offsets.data_ptr()), _P(vals.data_ptr()),
            _P(ids.data_ptr()), _P(se.data_ptr()), _P(sx.data_ptr()), _stream()))
        return vals, ids[:total], se[:total], sx[:total], offsets

    def adjoint_tensor(self, cloud: PointCloud, grad_out: torch.Tensor,
                       reduction=Sum, counts: torch.Tensor | None = None) -> torch.Tensor:
        """Gradient with respect to the cloud's fields (n, nfields), filtered
        point order, of a loss whose gradient with respect to this projection's
        image is grad_out (ngrid, n_out) (vt_adjoint).  `counts`: the rays'
        segment counts from project_tensor, if kept."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        s, e = self._rays(cloud)
        g = device_f64(grad_out, cloud._device).reshape(-1)
        nout = self._nout(cloud, reduction)
        if g.numel() != self.ngrid * nout:
            raise ValueError(f"grad_out must have {self.ngrid} x {nout} values, got {g.numel()}")
        out = torch.empty((cloud.get_npart(), cloud.get_nfields()), dtype=torch.float64,
                          device=cloud._device)
        _lib.check(_lib.lib().vt_adjoint(cloud._handle(), _P(s.data_ptr()), _P(e.data_ptr()),
                                         self.ngrid, int(reduction),
                                         _P(counts.data_ptr()) if counts is not None else None,
                                         _P(g.data_ptr()), _P(out.data_ptr()), _stream()))
        return out

    def makeDetailedProjection(self, cloud: PointCloud, reduction=Sum):
        v, ids, se, sx, off = self.segments_tensor(cloud, reduction)
        self._data = v
        self._segments = (ids, se, sx, off)

    def returnSegments(self):
        if self._segments is None:
            raise RuntimeError("No detailed projection data available")
        return tuple(t.cpu().numpy() for t in self._segments)


class Ray:
    """One ray (Ray, include/ray.hpp; binding src/bindings.cpp:277-307): the
    walk through a cloud and its reductions, on the GPU.  Segments are
    (cell id, s_enter, s_exit, ds) with cell ids as filtered indices, as the
    reference's get_segments gives them."""

    def __init__(self, start, end):
        self._start = np.ascontiguousarray(start, dtype=np.float64).reshape(1, 3)
        self._end = np.ascontiguousarray(end, dtype=np.float64).reshape(1, 3)
        self._segments = []
        self._vals = {}

    def _trace(self, cloud: PointCloud, mode):
        v, ids, se, sx, _ = Projection(self._start, self._end).segments_tensor(cloud, mode)
        ids, se, sx = ids.cpu().numpy(), se.cpu().numpy(), sx.cpu().numpy()
        self._segments = [(int(i), float(a), float(b), float(b - a))
                          for i, a, b in zip(ids, se, sx)]
        return [float(x) for x in v.cpu().numpy()[0]]

    def walk(self, cloud: PointCloud):
        """Ray::walk (src/ray.cpp:187-430): traces the ray; returns its
        segments (the reference's binding cannot convert its return value)."""
        self._trace(cloud, Sum)
        return self.get_segments()

    def integrate(self, cloud: PointCloud, mode=Sum) -> None:
        """Ray::integrate (src/ray.cpp:432-442): walk, then reduce by `mode`."""
        self._vals[int(mode)] = self._trace(cloud, ReductionMode(int(mode)))

    def get_dens_col(self) -> float:
        col = self._vals.get(int(Sum), [])
        return col[0] if col else 0.0

    def get_col(self):
        return list(self._vals.get(int(Sum), []))

    def get_max_val(self):
        return list(self._vals.get(int(Max), []))

    def get_min_val(self):
        return list(self._vals.get(int(Min), []))

    def get_vol_render_val(self):
        return list(self._vals.get(int(VolumeRender), []))

    def get_segments(self):
        return list(self._segments)


class Slice:
    """Slice (src/slice.hpp, binding src/bindings.cpp:259-275): nearest-cell
    field values on an nx x ny grid at a fixed depth."""

    def __init__(self, npix, extent, depth):
        self._npix = (int(npix[0]), int(npix[1]))
        self._extent = np.ascontiguousarray(np.asarray(extent, dtype=np.float64).ravel())
        if self._extent.size != 4:
            raise ValueError("extent must have 4 elements")
        self._depth = float(depth)
        self._data = None

    def slice_tensor(self, cloud: PointCloud) -> torch.Tensor:
        """makeSlice with the result on the device ((nx * ny), nfields)."""
        if not cloud.get_tree_built():
            raise RuntimeError("There is currently no valid tree for this point cloud")
        nx, ny = self._npix
        out = torch.zeros((nx * ny, cloud.get_nfields()), dtype=torch.float64,
                          device=cloud._device)
        _lib.check(_lib.lib().vt_slice(cloud._handle(), _ptr(self._extent), nx, ny,
                                       C.c_double(self._depth), _P(out.data_ptr()),
                                       _stream()))
        return out

    def makeSlice(self, cloud: PointCloud) -> None:
        self._data = self.slice_tensor(cloud)

    def returnSlice(self) -> np.ndarray:
        if self._data is None:
            raise RuntimeError("Slice has not yet been made")
        return self._data.cpu().numpy()
