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
