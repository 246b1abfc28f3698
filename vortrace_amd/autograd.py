"""Differentiable projections: gradients with respect to the cell fields.

``project(pc, fields, pos_start, pos_end, reduction)`` projects the rays
(pos_start, pos_end) through ``pc``'s mesh with ``fields`` -- a tensor in the
input point order of ``pc``, shape (N,) or (N, nfields) -- and is
differentiable with respect to ``fields``.  The derivative is that of the
reference's reduce_sum (src/reduction.cpp:8-21: a segment-length-weighted
scatter, for 'sum'/'integrate') and reduce_volume_render
(src/reduction.cpp:55-83: reverse compositing per ray, for 'volume').
Per-cell gradients are summed in ray-major segment order, so they are
deterministic.

The rays are walked once per forward/backward pair: the forward pass records
every segment (its cell, ray, place along the ray and length) in a device
pool while it projects, and the backward pass reorders those records instead
of walking the rays again (``vt_project_recorded`` / ``vt_adjoint_recorded``).
The pool is sized from the segment count of the previous forward pass on the
same cloud; on the first pass, or when the pool overflows, the backward pass
walks the rays for their segments (``vt_adjoint``).  Both give the same bits.

From the second pass on, when the rays are one to eight times the lanes the
kernel keeps resident, the forward pass also takes them longest first (by the
previous pass's segment counts, in whole runs of 64 neighbouring rays within
blocks of each ray partition), which shortens the kernel's tail; the order
never changes a result.

The reference has no derivatives; the tree and the segments are the
reference's, so the gradient is the exact derivative of its projection.
Points that the cloud filtered out (non-periodic padding filter,
src/pointcloud.cpp:55-108) get a zero gradient.
"""
from __future__ import annotations

import functools

import torch

from . import _lib
from .core import SEGREC_BYTES, Projection, ReductionMode

# Pool headroom over the previous pass's segment count: a fraction, and whole
# record chunks left partly filled by the kernel's lanes (at most one chunk of
# 32 records per lane of each of the two launches).
_POOL_GROWTH = 1.05
_POOL_SLACK = 2 * 131072 * 32


def _pool_capacity(estimate: int) -> int:
    return int(estimate * _POOL_GROWTH) + _POOL_SLACK


# Rays per block of the longest-first order: 2^15 rays of an image are a band
# of rows small enough that their cells stay in one XCD's L2.
_ORDER_BLOCK = 1 << 15


def _order_pays(n: int, device) -> bool:
    """Whether the longest-first order is worth its loss of ray locality: only
    when each resident lane takes more than one ray and at most eight (about
    512 lanes per CU are resident).  Measured on the forward kernel: 2 rays
    per lane (512^2 rays, 2M clustered cells) -9 %; 8 rays per lane (1024^2,
    volume) -2.5 %; half a ray per lane (256^2) +17 %; 32 rays per lane (16
    views of 512^2) +1.5 %."""
    lanes = 512 * torch.cuda.get_device_properties(device).multi_processor_count
    return lanes < n <= 8 * lanes


def tiled_slots(n: int, row: int, device=None) -> torch.Tensor:
    """The ray each fetch slot takes under the 8x8-tile order of images with
    `row` rays per row (vt_cloud_set_ray_rows; the identity for row 0 and
    for the rays past the last whole band of 8 rows)."""
    j = torch.arange(n, device=device, dtype=torch.int64)
    band = 8 * row
    nfull = (n // band) * band if row > 0 else 0
    if nfull == 0:
        return j
    jf = j[:nfull]
    b = jf // band
    q = jf - b * band
    t = q // 64
    cw = torch.clamp(row - 8 * t, max=8)
    u = q - 64 * t
    r = u // cw
    out = j.clone()
    out[:nfull] = (8 * b + r) * row + 8 * t + (u - r * cw)
    return out


@functools.lru_cache(maxsize=4)
def order_plan(n: int, row: int, parts: int, block: int, device: torch.device):
    """The parts of longest_first that depend only on the sizes: the fetch
    order (tiled_slots), each slot's run, the runs' first slots and lengths,
    and each run's block (its sort group)."""
    rays = tiled_slots(n, row, device)
    slot = torch.arange(n, device=device, dtype=torch.int64)
    psize = -(-n // parts)
    per_part = -(-psize // block)
    per_block = -(-block // 64)
    nrun = parts * per_part * per_block
    part = slot // psize
    pl = slot - part * psize
    run = (part * per_part + pl // block) * per_block + (pl % block) // 64
    # Run ids grow with the slot, so a run's slots start at the exclusive sum
    # of the earlier runs' lengths.
    run_len = torch.bincount(run, minlength=nrun)
    run_start = torch.cumsum(run_len, 0) - run_len
    group = torch.arange(nrun, device=device, dtype=torch.int64) // per_block
    return rays, slot, run, run_len, run_start, group


def longest_first(counts: torch.Tensor, row: int = 0,
                  block: int = _ORDER_BLOCK) -> torch.Tensor | None:
    """A ray order (int32 permutation of [0, n), for set_ray_order) from a
    previous pass's segment counts.  The rays are first put in fetch order
    (8x8 tiles of `row`-ray image rows, or index order for row 0); each ray
    partition (the kernels' XCD partitions are contiguous ranges of slots) is
    cut in blocks of `block` slots and each block in runs of 64 -- what one
    wave takes at once; a block takes its runs whole, longest first (by the
    run's largest count; slot order among equals).  The
    rays a lane takes last are then short ones, and a wave's rays are still
    neighbours.  None for 2^31 rays or more."""
    from .tuning import get_tuning
    n = counts.numel()
    if n == 0 or n >= 2 ** 31:
        return None
    dev = counts.device
    rays, slot, run, run_len, run_start, group = order_plan(
        n, int(row), get_tuning()["ray_parts"], int(block), dev)
    run_max = torch.zeros(run_len.numel(), dtype=torch.int64, device=dev)
    run_max.scatter_reduce_(0, run, counts.to(torch.int64)[rays], reduce="amax",
                            include_self=False)
    # Sort the runs (not the rays) and lay them out again in that order.
    span = 1 << 40  # above any segment count
    perm = torch.argsort(group * span + (span - 1 - run_max), stable=True)
    plen = run_len[perm]
    out_start = torch.cumsum(plen, 0) - plen
    at = torch.repeat_interleave(perm, plen, output_size=n)
    within = slot - torch.repeat_interleave(out_start, plen, output_size=n)
    return rays[run_start[at] + within].to(torch.int32)


class ProjectFields(torch.autograd.Function):
    """projection(fields) with the adjoint as its backward."""

    @staticmethod
    def forward(ctx, fields, pc, pos_start, pos_end, mode):
        pc.set_fields(fields)
        proj = Projection(pos_start, pos_end)
        dev = pc._cloud._device
        counts = torch.empty(proj.ngrid, dtype=torch.int64, device=dev)
        records, n_records = None, -1
        estimate = getattr(pc, "_segment_estimate", None)
        # The previous pass's counts order this pass's rays longest first
        # (speed only; any permutation gives the same bits).
        order = getattr(pc, "_longest_first", None)
        ordered = (order is not None and order.numel() == proj.ngrid
                   and not pc._cloud.get_ray_order_size())
        if ordered:
            pc._cloud.set_ray_order(order)
        try:
            if ctx.needs_input_grad[0] and estimate is not None:
                records = torch.empty((_pool_capacity(estimate), SEGREC_BYTES),
                                      dtype=torch.uint8, device=dev)
                out, n_records = proj.project_recorded_tensor(pc._cloud, mode, counts,
                                                              records)
                if n_records < 0:
                    records = None  # pool too small: the backward walks the rays
            else:
                out = proj.project_tensor(pc._cloud, mode, counts=counts)
            pc._forward_stats = _lib.last_call_stats()
        finally:
            if ordered:
                pc._cloud.set_ray_order(None)
        if ctx.needs_input_grad[0]:
            pc._segment_estimate = int(counts.sum().item())
            pc._longest_first = (longest_first(counts, pc._cloud.get_ray_rows())
                                 if _order_pays(counts.numel(), dev) else None)
        ctx.save_for_backward(fields, counts)
        ctx.records, ctx.n_records = records, n_records
        ctx.pc, ctx.proj, ctx.mode = pc, proj, mode
        return out[:, 0] if out.shape[1] == 1 else out

    @staticmethod
    def backward(ctx, grad):
        fields, counts = ctx.saved_tensors
        pc, mode = ctx.pc, ctx.mode
        if mode == ReductionMode.VolumeRender:
            pc.set_fields(fields)  # the volume adjoint reads the fields
        g = grad.reshape(grad.shape[0], -1).contiguous()
        if ctx.records is not None:
            gf = ctx.proj.adjoint_recorded_tensor(pc._cloud, g, mode, counts, ctx.records,
                                                  ctx.n_records)
        else:
            gf = ctx.proj.adjoint_tensor(pc._cloud, g, mode, counts=counts)
        # Statistics are per thread, and autograd may run this on its own.
        pc._adjoint_stats = _lib.last_call_stats()
        ctx.records = None
        if pc._keeps_all_in_order():
            full = gf
        else:
            full = torch.zeros((fields.shape[0], gf.shape[1]), dtype=torch.float64,
                               device=gf.device)
            full[pc._orig_ids_tensor(gf.device)] = gf
        return full.reshape(fields.shape).to(fields.dtype), None, None, None, None


def project(pc, fields: torch.Tensor, pos_start, pos_end,
            reduction: str = "integrate") -> torch.Tensor:
    """Differentiable projection of `fields` (see the module docstring).
    Returns (nrays,) for one output field, else (nrays, n_out)."""
    # RGBA validity is checked on the new fields, by the projection itself.
    mode = pc._validate_reduction(reduction, check_rgba=False)
    return ProjectFields.apply(fields, pc, pos_start, pos_end, mode)
