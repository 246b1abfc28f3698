"""Ray orders for the plain forward projection.

The ray kernels' time has a tail: the last rays handed out finish long after
the rest when they are long ones.  A forward pass of a fitting loop orders its
rays from the previous pass's segment counts (``autograd.longest_first``); a
plain projection has no counts before it runs, so it orders them by an
estimate: the cells a ray crosses grow with the cube root of the point
density along it, so each ray's cost is estimated as its length times the
mean of (points per cell)^(1/3) of a coarse G^3 histogram of the points,
sampled at S points along the ray (one ray per run of 64 neighbours).  The order only decides which lane walks
which ray when; every output keeps its ray's index and its bits.
"""
from __future__ import annotations

import numpy as np
import torch

_G = 32        # histogram cells per axis
_SAMPLES = 32  # samples along each ray


def density_grid(pc, device) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(points per cell)^(1/3) on a G^3 grid over the cloud's bounding box,
    as a flat float32 device tensor, with the box's low corner and size."""
    key = ("_cost_grid", str(device))
    cached = getattr(pc, "_cost_grid", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    bb = np.asarray(pc.boundbox, dtype=np.float64).reshape(3, 2)
    lo, size = bb[:, 0], np.maximum(bb[:, 1] - bb[:, 0], 1e-300)
    p = pc.pos_orig[pc.orig_ids] if len(pc.orig_ids) != len(pc.pos_orig) else pc.pos_orig
    ijk = np.clip(((p - lo) / size * _G).astype(np.int64), 0, _G - 1)
    flat = (ijk[:, 0] * _G + ijk[:, 1]) * _G + ijk[:, 2]
    h = np.bincount(flat, minlength=_G ** 3).astype(np.float32) ** (1.0 / 3.0)
    out = (torch.from_numpy(h).to(device),
           torch.tensor(lo, dtype=torch.float32, device=device),
           torch.tensor(size, dtype=torch.float32, device=device))
    pc._cost_grid = (key, out)
    return out


def ray_costs(pc, start: torch.Tensor, end: torch.Tensor) -> torch.Tensor:
    """Estimated cost of each ray (float32): its length (in box diagonals)
    times the sum of (points per cell)^(1/3) at S points along it."""
    grid, lo, size = density_grid(pc, start.device)
    s = start.float()
    d = end.float() - s
    t = (torch.arange(_SAMPLES, device=start.device, dtype=torch.float32) + 0.5) / _SAMPLES
    p = s[:, None, :] + t[None, :, None] * d[:, None, :]
    ijk = torch.floor((p - lo) / size * _G).long()
    ijk = torch.remainder(ijk, _G) if pc.periodic else ijk.clamp_(0, _G - 1)
    flat = (ijk[..., 0] * _G + ijk[..., 1]) * _G + ijk[..., 2]
    # length in units of the box diagonal: the cost does not scale with the
    # caller's coordinate units
    return grid[flat].sum(dim=1) * (torch.linalg.vector_norm(d, dim=1)
                                    / torch.linalg.vector_norm(size))


def estimated_order(pc, start: torch.Tensor, end: torch.Tensor, row: int = 0):
    """autograd.longest_first of estimated ray costs: an int32 permutation
    of the rays for set_ray_order.  The order moves whole runs of 64 rays in
    fetch order (an 8x8 tile of an image), whose rays share histogram cells,
    so only the middle ray of each run of 64 fetch slots is estimated and
    stands for the run."""
    from .autograd import _ORDER_BLOCK, longest_first, order_plan
    from .tuning import get_tuning
    n = start.shape[0]
    dev = start.device
    rays = order_plan(n, int(row), get_tuning()["ray_parts"], _ORDER_BLOCK, dev)[0]
    probe = rays[torch.clamp(torch.arange(0, n, 64, device=dev) + 32, max=n - 1)]
    cost = (ray_costs(pc, start[probe], end[probe]) * 1024.0).long()
    counts = torch.empty(n, dtype=torch.int64, device=dev)
    counts[rays] = torch.repeat_interleave(cost, 64)[:n]
    return longest_first(counts, row)
