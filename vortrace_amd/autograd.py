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

The reference has no derivatives; the tree and the segments are the
reference's, so the gradient is the exact derivative of its projection.
Points that the cloud filtered out (non-periodic padding filter,
src/pointcloud.cpp:55-108) get a zero gradient.
"""
from __future__ import annotations

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
        if ctx.needs_input_grad[0] and estimate is not None:
            records = torch.empty((_pool_capacity(estimate), SEGREC_BYTES),
                                  dtype=torch.uint8, device=dev)
            out, n_records = proj.project_recorded_tensor(pc._cloud, mode, counts, records)
            if n_records < 0:
                records = None  # pool too small: the backward walks the rays
        else:
            out = proj.project_tensor(pc._cloud, mode, counts=counts)
        pc._forward_stats = _lib.last_call_stats()
        if ctx.needs_input_grad[0]:
            pc._segment_estimate = int(counts.sum().item())
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
