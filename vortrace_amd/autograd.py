"""Differentiable projections: gradients with respect to the cell fields.

``project(pc, fields, pos_start, pos_end, reduction)`` projects the rays
(pos_start, pos_end) through ``pc``'s mesh with ``fields`` -- a tensor in the
input point order of ``pc``, shape (N,) or (N, nfields) -- and is
differentiable with respect to ``fields``.  The forward pass is the projection
kernel, which also records each ray's segment count; the backward pass is the
adjoint (``vt_adjoint``), which walks the rays once more for their segments:
segment-length-weighted scatter for 'sum'/'integrate' (the derivative of the
reference's reduce_sum, src/reduction.cpp:8-21) and reverse compositing per
ray for 'volume' (reduce_volume_render, src/reduction.cpp:55-83).  Per-cell
gradients are summed in ray-major segment order, so they are deterministic.

The reference has no derivatives; the tree and the segments are the
reference's, so the gradient is the exact derivative of its projection.
Points that the cloud filtered out (non-periodic padding filter,
src/pointcloud.cpp:55-108) get a zero gradient.
"""
from __future__ import annotations

import torch

from .core import Projection, ReductionMode


class ProjectFields(torch.autograd.Function):
    """projection(fields) with the adjoint as its backward."""

    @staticmethod
    def forward(ctx, fields, pc, pos_start, pos_end, mode):
        pc.set_fields(fields)
        proj = Projection(pos_start, pos_end)
        counts = torch.empty(proj.ngrid, dtype=torch.int64, device=pc._cloud._device)
        out = proj.project_tensor(pc._cloud, mode, counts=counts)
        ctx.save_for_backward(fields, counts)
        ctx.pc, ctx.proj, ctx.mode = pc, proj, mode
        return out[:, 0] if out.shape[1] == 1 else out

    @staticmethod
    def backward(ctx, grad):
        fields, counts = ctx.saved_tensors
        pc, mode = ctx.pc, ctx.mode
        if mode == ReductionMode.VolumeRender:
            pc.set_fields(fields)  # the volume adjoint reads the fields
        g = grad.reshape(grad.shape[0], -1).contiguous()
        gf = ctx.proj.adjoint_tensor(pc._cloud, g, mode, counts=counts)
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
