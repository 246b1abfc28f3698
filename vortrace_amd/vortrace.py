"""``ProjectionCloud``: the user-facing drop-in for vortrace's Python API.

Same constructor, methods, argument names, return shapes and exception
messages as the reference's vortrace/vortrace.py:24-338, with the ray walk
running on the GPU.  Differences that are supersets: ``center`` of
``grid_projection`` defaults to None (the reference requires it positionally,
vortrace.py:141-144, while its README calls it without); the ``*_tensor``
methods keep inputs and results on the device; ``project_autograd`` returns a
differentiable tensor.
"""
from __future__ import annotations

import contextlib
import logging

import numpy as np
import torch
from numpy.typing import ArrayLike

from . import grid as gr
from .core import PointCloud, Projection, ReductionMode, Slice

_log = logging.getLogger("vortrace")


class ProjectionCloud:
    """Projections through the Voronoi mesh of a point cloud."""

    _REDUCTION_MAP = {
        "integrate": ReductionMode.Sum,
        "sum": ReductionMode.Sum,
        "max": ReductionMode.Max,
        "min": ReductionMode.Min,
        "volume": ReductionMode.VolumeRender,
    }

    def __init__(self, pos: ArrayLike, fields: ArrayLike,
                 boundbox: list[float] | None = None,
                 vol: ArrayLike | None = None, *,
                 periodic: bool = False,
                 _skip_build_tree: bool = False) -> None:
        """vortrace.py:40-104: FP64 copies, field count, default bounding box
        from the data, load (filter / validate), build the device tree."""
        self.pos_orig = np.ascontiguousarray(pos, dtype=np.float64)
        self.fields_orig = np.ascontiguousarray(fields, dtype=np.float64)
        self.vol_orig = (np.ascontiguousarray(vol, dtype=np.float64)
                         if vol is not None else None)
        self.periodic = periodic
        if self.fields_orig.ndim == 1:
            self._nfields = 1
        elif self.fields_orig.ndim == 2:
            self._nfields = self.fields_orig.shape[1]
        else:
            raise ValueError("fields must be 1D (npart,) or 2D (npart, nfields)")
        if boundbox is None:
            p = self.pos_orig
            boundbox = [p[:, 0].min(), p[:, 0].max(), p[:, 1].min(),
                        p[:, 1].max(), p[:, 2].min(), p[:, 2].max()]
        self.boundbox = boundbox
        self._cloud = PointCloud()
        self._cloud.loadPoints(self.pos_orig, self.fields_orig, boundbox,
                               self.vol_orig, periodic)
        if not _skip_build_tree:
            self._cloud.buildTree()
        self.orig_ids = self._cloud.get_orig_ids()
        _log.info("Applied bounding box filter: %d -> %d particles",
                  len(self.pos_orig), len(self.orig_ids))

    # ------------------------------------------------------------------
    def _validate_reduction(self, reduction: str, check_rgba: bool = True) -> ReductionMode:
        """vortrace.py:106-125 (same messages)."""
        mode = self._REDUCTION_MAP.get(reduction)
        if mode is None:
            raise ValueError(f"Unknown reduction {reduction!r}. "
                             f"Use one of {list(self._REDUCTION_MAP)}")
        if mode == ReductionMode.VolumeRender:
            if self._nfields != 4:
                raise ValueError("Volume rendering requires exactly 4 fields "
                                 f"(R, G, B, alpha), got {self._nfields}")
            if check_rgba and not self._cloud.get_valid_rgba():
                raise ValueError("Volume rendering requires R, G, B values in "
                                 "[0, 1] and alpha >= 0")
        return mode

    def _out_fields(self, mode) -> int:
        return 3 if mode == ReductionMode.VolumeRender else self._nfields

    def _require_tree(self):
        if not self._cloud.get_tree_built():
            raise RuntimeError(
                "There is currently no valid tree for this point cloud")

    # ------------------------------------------------------------------
    def grid_projection(self, extent: ArrayLike, nres, bounds: ArrayLike,
                        center: ArrayLike | None = None, *,
                        proj: str | None = None, yaw: float = 0.,
                        pitch: float = 0., roll: float = 0.,
                        reduction: str = "integrate") -> np.ndarray:
        """vortrace.py:141-195: (nx, ny), (nx, ny, nfields) or (nx, ny, 3)."""
        mode = self._validate_reduction(reduction)
        start, end = gr.generate_projection_grid(extent, nres, bounds, center,
                                                 proj=proj, yaw=yaw,
                                                 pitch=pitch, roll=roll)
        shape = start.shape
        p = Projection(start.reshape(-1, 3), end.reshape(-1, 3))
        self._require_tree()
        with self._image_rows(shape[1]):
            if not self._estimated_order(p, mode):
                p.makeProjection(self._cloud, mode)
            else:
                try:
                    p.makeProjection(self._cloud, mode)
                finally:
                    self._cloud.set_ray_order(None)
        dat = p.returnProjection()
        nout = self._out_fields(mode)
        if nout == 1:
            return np.reshape(dat, shape[:-1])
        return np.reshape(dat, (*shape[:-1], nout))

    def projection(self, pos_start: ArrayLike, pos_end: ArrayLike, *,
                   reduction: str = "integrate") -> np.ndarray:
        """vortrace.py:197-227: (N,) or (N, n_out)."""
        mode = self._validate_reduction(reduction)
        s = np.ascontiguousarray(pos_start, dtype=np.float64)
        e = np.ascontiguousarray(pos_end, dtype=np.float64)
        if (s.ndim != 2 or e.ndim != 2 or s.shape != e.shape
                or s.shape[1] != 3):
            raise ValueError("pos_start / pos_end must be 2D arrays of "
                             "identical shape (N,3)")
        p = Projection(s, e)
        self._require_tree()
        p.makeProjection(self._cloud, mode)
        return p.returnProjection()

    def traced_projection(self, pos_start: ArrayLike, pos_end: ArrayLike,
                          return_midpoint: bool = True, *,
                          reduction: str = "integrate",
                          flatten: bool = False) -> tuple:
        """vortrace.py:229-338: values plus per-segment cell ids, s and ds,
        for one ray or a batch (list or flat-with-offsets format)."""
        s = np.asarray(pos_start)
        e = np.asarray(pos_end)
        single = False
        if s.ndim == 1 and e.ndim == 1:
            if s.shape != (3,) or e.shape != (3,):
                raise ValueError("If 1D, pos_start and pos_end must both "
                                 "have shape (3,)")
            s, e = s[np.newaxis, :], e[np.newaxis, :]
            single = True
        elif s.ndim == 2 and e.ndim == 2:
            if s.shape != e.shape or s.shape[1] != 3:
                raise ValueError("pos_start and pos_end must be 2D arrays "
                                 "of identical shape (N, 3)")
            single = s.shape[0] == 1
        else:
            raise ValueError("pos_start and pos_end must both be 1D shape "
                             "(3,) or both 2D shape (N, 3)")
        s = np.ascontiguousarray(s, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        mode = self._validate_reduction(reduction)
        p = Projection(s, e)
        self._require_tree()
        p.makeDetailedProjection(self._cloud, mode)
        values = p.returnProjection()
        cell_ids, s_enter, s_exit, offsets = p.returnSegments()
        cell_ids = self.orig_ids[cell_ids]
        ds_vals = s_exit - s_enter
        s_vals = (s_enter + s_exit) / 2.0 if return_midpoint else s_enter
        if single:
            if mode == ReductionMode.VolumeRender or self._nfields > 1:
                val = values[0]
            else:
                val = values[0] if values.ndim == 1 else values[0, 0]
            return val, cell_ids, s_vals, ds_vals
        if flatten:
            return values, cell_ids, s_vals, ds_vals, offsets
        idx = range(len(offsets) - 1)
        return (values,
                [cell_ids[offsets[i]:offsets[i + 1]] for i in idx],
                [s_vals[offsets[i]:offsets[i + 1]] for i in idx],
                [ds_vals[offsets[i]:offsets[i + 1]] for i in idx])

    def single_projection(self, *args, **kwargs):
        """Removed in the reference (vortrace.py:340-345)."""
        raise AttributeError("'single_projection' has been removed. "
                             "Use 'traced_projection' instead.")

    # -- device-resident API ------------------------------------------------
    def set_ray_rows(self, row: int) -> None:
        """Tells the ray kernels that the rays of this cloud's coming calls
        are the pixels of row-major images with `row` rays per row (0, the
        default: no image order); they then take them in 8x8-pixel tiles.
        Speed only: results keep their bits and their order."""
        self._require_tree()
        self._cloud.set_ray_rows(row)

    def set_ray_order(self, order: torch.Tensor | None) -> None:
        """Sets the order in which the ray kernels take the rays of this
        cloud's coming calls with len(order) rays: an int32 permutation of
        [0, n) on the device, e.g. vortrace_amd.autograd.longest_first(counts)
        from a previous pass's segment counts; None clears it.  Speed only."""
        self._require_tree()
        self._cloud.set_ray_order(order)

    @contextlib.contextmanager
    def _image_rows(self, row: int):
        prev = self._cloud.get_ray_rows()
        self._cloud.set_ray_rows(row)
        try:
            yield
        finally:
            self._cloud.set_ray_rows(prev)

    def projection_tensor(self, pos_start: torch.Tensor, pos_end: torch.Tensor,
                          *, reduction: str = "integrate") -> torch.Tensor:
        """Like ``projection`` but inputs and output stay on the device;
        returns (N, n_out) FP64."""
        mode = self._validate_reduction(reduction)
        self._require_tree()
        proj = Projection(pos_start, pos_end)
        if not self._estimated_order(proj, mode):
            return proj.project_tensor(self._cloud, mode)
        try:
            return proj.project_tensor(self._cloud, mode)
        finally:
            self._cloud.set_ray_order(None)

    def _estimated_order(self, proj, mode=ReductionMode.Sum) -> bool:
        """Sets a longest-first order of the rays by estimated cost
        (PointCloud.order_rays, built on the device) when it pays
        (autograd._order_pays, with the lanes of the kernel the call
        launches) and the caller set no order of its own; returns whether it
        did (the caller then clears it after its call)."""
        from .autograd import _order_pays
        if self._cloud.get_ray_order_size() or not _order_pays(proj.ngrid,
                                                               self._cloud.lanes(mode)):
            return False
        s, e = proj._rays(self._cloud)
        self._cloud.order_rays(s, e, row=self._cloud.get_ray_rows())
        return True

    def batch_projection_tensor(self, views, *, reduction: str = "integrate"):
        """Projects several ray sets -- views of any sizes and directions,
        given as (pos_start, pos_end) pairs -- in ONE launch of the ray kernel:
        every ray joins one global work queue, which evens out the very
        different crossings per ray of the views.  Returns one device tensor
        (n_i, n_out) per view."""
        mode = self._validate_reduction(reduction)
        self._require_tree()
        dev = self._cloud._device
        starts, ends, sizes = [], [], []
        for ps, pe in views:
            s = torch.as_tensor(ps, dtype=torch.float64).reshape(-1, 3)
            e = torch.as_tensor(pe, dtype=torch.float64).reshape(-1, 3)
            if s.shape != e.shape:
                raise ValueError("pos_start and pos_end must have the same shape")
            starts.append(s.to(dev))
            ends.append(e.to(dev))
            sizes.append(s.shape[0])
        if not sizes:
            return []
        out = Projection(torch.cat(starts), torch.cat(ends)).project_tensor(self._cloud, mode)
        return list(torch.split(out, sizes))

    def batch_grid_projection(self, views, *, reduction: str = "integrate"):
        """grid_projection for many views in ONE launch, with the rays built
        on the device (grid.generate_projection_grids_device) from each
        view's grid_projection arguments (a list of dicts: extent, nres,
        bounds, center, proj, yaw, pitch, roll).  Returns one device tensor
        per view shaped like grid_projection's result."""
        mode = self._validate_reduction(reduction)
        self._require_tree()
        dev = self._cloud._device
        start, end, sizes = gr.generate_projection_grids_device(views, dev)
        if not sizes:
            return []
        shapes = [(n, n) if isinstance(n, (int, np.integer)) else tuple(n)
                  for n in (v["nres"] for v in views)]
        rows = {int(ny) for _, ny in shapes}
        with self._image_rows(rows.pop() if len(rows) == 1 else 0):
            out = Projection(start, end).project_tensor(self._cloud, mode)
        nout = self._out_fields(mode)
        res = []
        for (nx, ny), o in zip(shapes, torch.split(out, sizes)):
            res.append(o.reshape(int(nx), int(ny)) if nout == 1 else
                       o.reshape(int(nx), int(ny), nout))
        return res

    def batch_projection(self, views, *, reduction: str = "integrate"):
        """batch_projection_tensor with numpy results shaped like
        ``projection``'s: (n_i,) for one output field, else (n_i, n_out)."""
        return [o.cpu().numpy()[:, 0].copy() if o.shape[1] == 1 else o.cpu().numpy()
                for o in self.batch_projection_tensor(views, reduction=reduction)]

    def segments_tensor(self, pos_start, pos_end, *,
                        reduction: str = "integrate"):
        """Device-resident traced projection: (values, cell_ids (original
        input rows), s_enter, s_exit, offsets)."""
        mode = self._validate_reduction(reduction)
        self._require_tree()
        v, ids, se, sx, off = Projection(pos_start, pos_end).segments_tensor(
            self._cloud, mode)
        oid = torch.from_numpy(self.orig_ids).to(ids.device)
        return v, oid[ids], se, sx, off

    # -- fields on the device and gradients ---------------------------------
    def _orig_ids_tensor(self, device) -> torch.Tensor:
        t = getattr(self, "_oid_t", None)
        if t is None or t.device != device:
            t = torch.from_numpy(self.orig_ids).to(device)
            self._oid_t = t
        return t

    def set_fields(self, fields) -> None:
        """Replaces the fields ((N,) or (N, nfields), input point order, any
        device) without rebuilding the tree."""
        self._require_tree()
        dev = self._cloud._device
        f = fields.detach() if isinstance(fields, torch.Tensor) else torch.as_tensor(fields)
        f = f.to(device=dev, dtype=torch.float64).reshape(len(self.pos_orig), -1)
        if f.shape[1] != self._nfields:
            raise ValueError(f"fields must have {self._nfields} columns, got {f.shape[1]}")
        if self._keeps_all_in_order():
            self._cloud.set_fields_tensor(f.contiguous())
        else:
            self._cloud.set_fields_tensor(f[self._orig_ids_tensor(dev)].contiguous())

    def _keeps_all_in_order(self) -> bool:
        """True when the filtered points are all input points in input order
        (no padding filter took any out), so no reordering is needed."""
        flag = getattr(self, "_ids_identity", None)
        if flag is None:
            ids = self.orig_ids
            flag = len(ids) == len(self.pos_orig) and bool(
                np.array_equal(ids, np.arange(len(ids), dtype=ids.dtype)))
            self._ids_identity = flag
        return flag

    def projection_grad(self, pos_start, pos_end, fields, *,
                        reduction: str = "integrate") -> torch.Tensor:
        """Projection of `fields` that is differentiable with respect to them
        (torch.autograd; the backward is the adjoint kernel)."""
        from .autograd import project
        return project(self, fields, pos_start, pos_end, reduction)

    # -- slice (vortrace.py:347-385) ------------------------------------------
    def slice(self, extent: ArrayLike, nres, depth: float) -> np.ndarray:
        """Nearest-cell field values on a pixel grid at z = depth (vortrace.py:347-385)."""
        npix = (nres, nres) if isinstance(nres, int) else tuple(nres)
        extent_arr = np.asarray(extent, dtype=np.float64).ravel()
        if extent_arr.size != 4:
            raise ValueError("extent must have 4 elements [xmin, xmax, ymin, ymax]")
        sl = Slice(list(npix), list(extent_arr), float(depth))
        sl.makeSlice(self._cloud)
        dat = sl.returnSlice()
        if self._nfields == 1:
            return dat.reshape(npix[0], npix[1])
        return dat.reshape(npix[0], npix[1], self._nfields)

    # -- tree serialisation (vortrace.py:391-401) -----------------------------
    def save_tree_bytes(self) -> bytes | None:
        """The kD-tree as nanoflann's index dump, or None if not built."""
        if self._cloud.get_tree_built():
            return self._cloud.saveTreeToBytes()
        return None

    def load_tree_bytes(self, data: bytes) -> None:
        """Restores the kD-tree from a nanoflann index dump."""
        self._cloud.loadTreeFromBytes(data)
