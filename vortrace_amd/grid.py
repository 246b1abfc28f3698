"""Ray-grid generation for grid projections.

The ray endpoints decide every later bit of the walk, so this module produces
exactly the FP64 arrays the reference's ``vortrace/grid.py`` produces: pixel
centres from ``numpy.linspace`` (grid.py:35-44), ``meshgrid(indexing='ij')``
so that ``[i, j]`` is (x_i, y_j), the six named Cartesian projections with a
handedness flip (grid.py:13-23, 163-172) and a Tait-Bryan rotation about a
centre applied with ``numpy.dot(grid, R.T)`` (grid.py:86-97, 176-190).  The
same numpy routines are used so the bits match by construction.
"""
from __future__ import annotations

import numpy as np
from numpy.typing import ArrayLike

# name -> (yaw, pitch, roll, handedness); reference grid.py:13-23.
_PLANES = {
    "xy": (0.0, 0.0, 0.0, 1),
    "yz": (np.pi / 2.0, 0.0, np.pi / 2.0, 1),
    "zx": (3.0 * np.pi / 2.0, 3.0 * np.pi / 2.0, 0.0, 1),
    "yx": (np.pi / 2.0, 0.0, np.pi, -1),
    "xz": (0.0, 0.0, np.pi / 2.0, -1),
    "zy": (0.0, 3.0 * np.pi / 2.0, 0.0, -1),
}


def _plane_angles(proj: str):
    try:
        return _PLANES[proj]
    except KeyError:
        raise ValueError(f"Unknown projection {proj!r}. "
                         f"Valid options: {list(_PLANES)}") from None


def _pixel_centres(lo, hi, n):
    d = (hi - lo) / n
    return np.linspace(lo + 0.5 * d, hi - 0.5 * d, n)


def generate_base_grid(extent: ArrayLike, nres) -> np.ndarray:
    """(nx, ny, 3) pixel-centre grid in the z=0 plane (reference grid.py:47-83)."""
    ext = np.array(extent)
    if ext.ravel().size == 2:
        ext = np.array([ext, ext])
    elif ext.ravel().size != 4:
        raise ValueError("extent must be either (2,2) or (2,) array")
    if isinstance(nres, int):
        nres = (nres, nres)
    xs = _pixel_centres(ext[0][0], ext[0][1], nres[0])
    ys = _pixel_centres(ext[1][0], ext[1][1], nres[1])
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    grid = np.zeros((nres[0], nres[1], 3))
    grid[..., 0] = gx
    grid[..., 1] = gy
    return grid


def rotation_matrix(yaw, pitch, roll) -> np.ndarray:
    """Tait-Bryan rotation matrix (reference grid.py:86-97)."""
    ca, sa = np.cos(yaw), np.sin(yaw)
    cb, sb = np.cos(pitch), np.sin(pitch)
    cg, sg = np.cos(roll), np.sin(roll)
    return np.array([
        [ca * cb, ca * sb * sg - sa * cg, ca * sb * cg + sa * sg],
        [sa * cb, sa * sb * sg + ca * cg, sa * sb * cg - ca * sg],
        [-sb, cb * sg, cb * cg],
    ])


def generate_projection_grid(extent: ArrayLike, nres, bounds: ArrayLike,
                             center: ArrayLike | None = None, *,
                             proj: str | None = None, yaw: float = 0.,
                             pitch: float = 0., roll: float = 0.):
    """Start and end points, each (nx, ny, 3), of a grid projection.

    Follows reference grid.py:100-192: z is set to ``bounds`` before any
    rotation; a named ``proj`` overrides the angles and flips ``bounds`` when
    its handedness disagrees with theirs; with ``center`` None no rotation is
    applied.
    """
    start = generate_base_grid(extent, nres)
    end = np.copy(start)
    if proj is not None:
        yaw, pitch, roll, hand = _plane_angles(proj)
    else:
        hand = np.sign(bounds[1] - bounds[0])
    if hand != np.sign(bounds[1] - bounds[0]):
        bounds = np.flip(bounds)
    start[..., 2] = bounds[0]
    end[..., 2] = bounds[1]
    if center is not None:
        start -= center
        end -= center
        rot = rotation_matrix(yaw, pitch, roll)
        start = np.dot(start, rot.T)
        end = np.dot(end, rot.T)
        start += center
        end += center
    return start, end


def view_params(extent: ArrayLike, nres, bounds: ArrayLike,
                center: ArrayLike | None = None, *, proj: str | None = None,
                yaw: float = 0., pitch: float = 0., roll: float = 0.):
    """The parameters of generate_projection_grid(...) as one C view record
    (vt_view) for the device generator: extent, resolution, the bounds after
    the handedness flip, and the rotation matrix (computed here, by numpy,
    exactly as generate_projection_grid computes it)."""
    from . import _lib
    ext = np.array(extent)
    if ext.ravel().size == 2:
        ext = np.array([ext, ext])
    elif ext.ravel().size != 4:
        raise ValueError("extent must be either (2,2) or (2,) array")
    if isinstance(nres, (int, np.integer)):
        nres = (int(nres), int(nres))
    if proj is not None:
        yaw, pitch, roll, hand = _plane_angles(proj)
    else:
        hand = np.sign(bounds[1] - bounds[0])
    if hand != np.sign(bounds[1] - bounds[0]):
        bounds = np.flip(bounds)
    v = _lib.View()
    for k, val in enumerate((ext[0][0], ext[0][1], ext[1][0], ext[1][1])):
        v.extent[k] = float(val)
    v.bounds[0], v.bounds[1] = float(bounds[0]), float(bounds[1])
    v.nx, v.ny = int(nres[0]), int(nres[1])
    if center is not None:
        c = np.zeros(3) + np.asarray(center)  # float64, as start -= center promotes
        rot = rotation_matrix(yaw, pitch, roll)
        for k in range(3):
            v.center[k] = float(c[k])
        for k in range(9):
            v.rot[k] = float(rot.flat[k])
        v.rotate = 1
    return v


def generate_projection_grids_device(views, device=None):
    """generate_projection_grid for many views at once on the device
    (vt_grid_rays).  `views` is a list of dicts of generate_projection_grid's
    arguments.  Returns (start, end, sizes): (R, 3) float64 device tensors
    holding every view's rays in turn (view v's ray i * ny + j is its pixel
    (i, j)) and the rays per view."""
    import ctypes as C
    import torch
    from . import _lib
    dev = torch.device(device) if device is not None else torch.device("cuda")
    recs = [view_params(**v) for v in views]
    sizes = [int(r.nx * r.ny) for r in recs]
    n = sum(sizes)
    start = torch.empty((n, 3), dtype=torch.float64, device=dev)
    end = torch.empty((n, 3), dtype=torch.float64, device=dev)
    arr = (_lib.View * max(1, len(recs)))(*recs)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.lib().vt_grid_rays(arr, len(recs), C.c_void_p(start.data_ptr()),
                                           C.c_void_p(end.data_ptr()), C.c_void_p(stream)))
    return start, end, sizes
