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
The pool is sized from the segment count of the previous forward pass over
the same rays (the same ray tensors or arrays, unchanged; a few ray sets are
remembered per cloud), or on the first pass over a ray set from a sample
(every k-th ray, about 4,096 of them, projected on its own); when the pool
overflows, the backward pass walks the rays for their segments
(``vt_adjoint``).  Both give the same bits.

From the second pass over a ray set on, when the rays are one to eight times
the lanes the kernel keeps resident, the forward pass also takes them longest
first (by the previous pass's segment counts for the same rays, in whole runs
of 64 neighbouring rays within blocks of each ray partition), which shortens
the kernel's tail; the order never changes a result.

The reference has no derivatives; the tree and the segments are the
reference's, so the gradient is the exact derivative of its projection.
Points that the cloud filtered out (non-periodic padding filter,
src/pointcloud.cpp:55-108) get a zero gradient.
"""
from __future__ import annotations

import collections
import functools
import warnings

import numpy as np
import torch

from . import _lib
from .core import SEGREC_BYTES, Projection, ReductionMode

# Pool headroom over the previous pass's segment count: a fraction, and the
# records the kernels' lanes can leave unused (at most one partly filled chunk
# of 32 records per resident lane, in each of the two launches of a call).
_POOL_GROWTH = 1.05
_REC_CHUNK = 32


def _pool_capacity(estimate: int, lanes: int) -> int:
    return int(estimate * _POOL_GROWTH) + 2 * lanes * _REC_CHUNK


# First pass over a ray set: segments estimated from every k-th ray (about
# _SAMPLE_RAYS of them), scaled up with headroom for the sampling error.
# Measured on 2M and 1M clustered meshes (512^2 and 256^2 z-rays): 2,048 to
# 4,096 sampled rays put the total within 0.5 % of the true one, and their
# projection takes 2-3 ms (65,536 of them took 21 ms).
_SAMPLE_RAYS = 4096
_SAMPLE_HEADROOM = 1.25


def _sampled_estimate(cloud, mode, pos_start, pos_end, n: int) -> int:
    """Segments of all n rays, estimated by projecting every k-th ray."""
    k = n // _SAMPLE_RAYS
    if k < 2:
        return -1  # too few rays to sample: the first pass walks twice
    s = torch.as_tensor(pos_start).reshape(-1, 3)[::k]
    e = torch.as_tensor(pos_end).reshape(-1, 3)[::k]
    p = Projection(s, e)
    c = torch.empty(p.ngrid, dtype=torch.int64, device=cloud._device)
    p.project_tensor(cloud, mode, counts=c)
    return int(float(c.sum().item()) * (n / p.ngrid) * _SAMPLE_HEADROOM)


# Rays per block of the longest-first order: 2^15 rays of an image are a band
# of rows small enough that their cells stay in one XCD's L2.
_ORDER_BLOCK = 1 << 15


def _order_pays(n: int, lanes: int) -> bool:
    """Whether the longest-first order is worth its loss of ray locality: only
    when each of the kernel's `lanes` resident lanes takes more than one ray
    and at most eight.  Measured on the forward kernel: 2 rays per lane (512^2
    rays, 2M clustered cells) -9 %; 8 rays per lane (1024^2, volume) -2.5 %;
    half a ray per lane (256^2) +17 %; 32 rays per lane (16 views of 512^2)
    +1.5 %."""
    return lanes < n <= 8 * lanes


def _rays_key(pos_start, pos_end):
    """Identity of a ray set between passes: tensors by storage, shape,
    version counter and device; arrays by address and shape.  Used only to
    size the record pool and order the rays (speed, never results)."""
    def one(x):
        if isinstance(x, torch.Tensor):
            return ("t", x.data_ptr(), tuple(x.shape), x._version, str(x.device), str(x.dtype))
        a = np.asarray(x)
        return ("n", a.__array_interface__["data"][0], a.shape, a.dtype.str)
    return one(pos_start), one(pos_end)


class _RayState:
    """What a forward pass leaves for the next pass over the same rays: the
    segment total (record pool size) and, when ordering pays, the per-ray
    segment counts the next pass orders its rays by."""
    __slots__ = ("estimate", "counts")

    def __init__(self, estimate, counts):
        self.estimate = estimate
        self.counts = counts


# Ray sets remembered per cloud: enough for a fitting loop over 16 views, each
# with its own ray tensors.  A state holds an int64 count per ray only when
# ordering pays (at most 8 rays per resident lane, 8 MB), so at most 128 MB.
_MAX_RAY_STATES = 16


def _ray_states(pc) -> collections.OrderedDict:
    st = getattr(pc, "_ray_state_lru", None)
    if st is None:
        st = collections.OrderedDict()
        pc._ray_state_lru = st
    return st


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
    # Sort the runs (not the rays) and lay them out again in that order; a
    # run's largest count is clamped below the span so groups stay apart.
    span = 1 << 40
    run_max.clamp_(0, span - 1)
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
        cloud = pc._cloud
        dev = cloud._device
        key = _rays_key(pos_start, pos_end)
        states = _ray_states(pc)
        state = states.get(key)
        if state is not None:
            states.move_to_end(key)
        counts = torch.empty(proj.ngrid, dtype=torch.int64, device=dev)
        records, n_records = None, -1
        # The previous pass's counts over the same rays order this pass's
        # rays longest first (speed only; any permutation gives the same bits).
        prev = state.counts if state is not None else None
        ordered = (prev is not None and prev.numel() == proj.ngrid
                   and not cloud.get_ray_order_size())
        if ordered:
            cloud.order_rays(counts=prev, row=cloud.get_ray_rows())
        try:
            estimate = -1
            if ctx.needs_input_grad[0]:
                estimate = (state.estimate if state is not None else
                            _sampled_estimate(cloud, mode, pos_start, pos_end, proj.ngrid))
            if estimate >= 0:
                lanes = cloud.lanes(mode, recorded=True)
                records = torch.empty((_pool_capacity(estimate, lanes), SEGREC_BYTES),
                                      dtype=torch.uint8, device=dev)
                out, n_records = proj.project_recorded_tensor(cloud, mode, counts, records)
                if n_records < 0:
                    records = None  # pool too small: the backward walks the rays
                    # (a sampled estimate may fall short; a previous pass's count
                    # over the same rays should not)
                    if state is not None and not getattr(pc, "_pool_warned", False):
                        pc._pool_warned = True
                        warnings.warn("vortrace_amd.autograd: the segment record pool "
                                      "overflowed; the backward pass walks the rays again "
                                      "(same result, slower)", RuntimeWarning, stacklevel=2)
            else:
                out = proj.project_tensor(cloud, mode, counts=counts)
            pc._forward_stats = _lib.last_call_stats()
        finally:
            if ordered:
                cloud.set_ray_order(None)
        if ctx.needs_input_grad[0]:
            lanes = cloud.lanes(mode, recorded=True)
            states[key] = _RayState(int(counts.sum().item()),
                                    counts if _order_pays(counts.numel(), lanes) else None)
            states.move_to_end(key)
            while len(states) > _MAX_RAY_STATES:
                states.popitem(last=False)
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
            pc._adjoint_path = "recorded"
        else:
            gf = ctx.proj.adjoint_tensor(pc._cloud, g, mode, counts=counts)
            pc._adjoint_path = "walk"
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
