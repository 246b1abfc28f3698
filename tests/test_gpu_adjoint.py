"""GPU tests of the adjoint (d image / d cell fields) and its autograd wrapper.

Checks, for the sum and volume reductions:
* Jacobian rows: for a one-hot image gradient the field gradient is the row
  J[r, c] = sum of ds over ray r's segments in cell c, built here from the
  oracle's segment lists (oracle/vt_oracle.c, pinned to the reference) in
  segment order -- bit-identical.
* Central finite differences of a squared loss through torch.autograd,
  relative error <= 1e-6 (the tolerance stated for gradients).
* Determinism: two backward passes give the same bits.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import grid, synthetic  # noqa: E402

FD_RTOL = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rays(npix, lo=0.0, hi=100.0):
    s, e = grid.generate_projection_grid([lo, hi], npix, [lo, hi], None)
    return s.reshape(-1, 3), e.reshape(-1, 3)


def _jacobian_rows(oc, s, e, rays, n):
    """J[r, :] over filtered indices from the oracle's segments."""
    _, ids, se, sx, off = oc.detailed(s, e, "sum")
    rows = np.zeros((len(rays), n))
    for k, r in enumerate(rays):
        for j in range(off[r], off[r + 1]):
            rows[k, ids[j]] += sx[j] - se[j]
    return rows


@pytest.mark.parametrize("periodic", [False, True])
def test_sum_adjoint_jacobian_rows(periodic):
    _say(f"sum adjoint Jacobian rows periodic={periodic}")
    pos, rho = synthetic.uniform_mesh(3000, seed=11)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    s, e = _rays(12)
    dev = torch.device("cuda")
    ds_, de_ = torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev)
    n = len(pc.orig_ids)
    rays = [0, 5, 17, 77, len(s) - 1]
    expect = _jacobian_rows(oc, s, e, rays, n)
    for k, r in enumerate(rays):
        g = torch.zeros(len(s), 1, dtype=torch.float64, device=dev)
        g[r, 0] = 1.0
        got = vt.Projection(ds_, de_).adjoint_tensor(pc._cloud, g, vt.ReductionMode.Sum)
        np.testing.assert_array_equal(got.cpu().numpy()[:, 0], expect[k],
                                      err_msg=f"Jacobian row of ray {r}")


@pytest.mark.parametrize("periodic", [False, True])
def test_sum_gradient_finite_differences(periodic):
    _say(f"sum gradient vs central differences periodic={periodic}")
    pos, rho = synthetic.uniform_mesh(4000, seed=12)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    s, e = _rays(16)
    dev = torch.device("cuda")
    ds_, de_ = torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev)
    rng = np.random.default_rng(3)
    rho_t = torch.from_numpy(rho).to(dev)
    target = pc.projection_tensor(ds_, de_)[:, 0] * torch.from_numpy(
        np.exp(0.1 * rng.standard_normal(len(s)))).to(dev)

    def loss(f):
        p = vt.autograd.project(pc, f, ds_, de_, "sum")
        return ((p - target) ** 2).sum()

    f = rho_t.clone().requires_grad_(True)
    L = loss(f)
    L.backward()
    grad = f.grad.detach().cpu().numpy()
    cells = np.flatnonzero(grad)[rng.permutation(np.count_nonzero(grad))[:64]]
    assert len(cells) >= 32
    for c in cells:
        h = 1e-4 * abs(rho[c]) + 1e-8
        fp, fm = rho_t.clone(), rho_t.clone()
        fp[c] += h
        fm[c] -= h
        fd = (loss(fp).item() - loss(fm).item()) / (2 * h)
        assert abs(fd - grad[c]) <= FD_RTOL * max(abs(fd), abs(grad[c])) + 1e-12 * abs(L.item()), \
            (c, fd, grad[c])


def test_volume_gradient_finite_differences():
    _say("volume gradient vs central differences")
    pos, rho = synthetic.uniform_mesh(3000, seed=13)
    bb = [0, 100, 0, 100, 0, 100]
    rng = np.random.default_rng(4)
    rgba = np.empty((len(pos), 4))
    rgba[:, :3] = 0.1 + 0.8 * rng.random((len(pos), 3))
    rgba[:, 3] = 0.02 + 0.05 * rng.random(len(pos))
    pc = vt.ProjectionCloud(pos, rgba, boundbox=bb, periodic=True)
    s, e = _rays(12)
    dev = torch.device("cuda")
    ds_, de_ = torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev)
    w = torch.from_numpy(rng.standard_normal((len(s), 3))).to(dev)
    f0 = torch.from_numpy(rgba).to(dev)

    def loss(f):
        return (vt.autograd.project(pc, f, ds_, de_, "volume") * w).sum()

    f = f0.clone().requires_grad_(True)
    L = loss(f)
    L.backward()
    grad = f.grad.detach().cpu().numpy()
    nz = np.argwhere(grad != 0)
    pick = nz[rng.permutation(len(nz))[:64]]
    assert len(pick) >= 32
    # Steps: the loss is linear in R, G, B (any step is exact up to rounding,
    # eps * |L| / h ~ 1e-12 here); in alpha the central-difference truncation
    # error is ~ (h * ds)^2 / 6 relative (< 1e-8 for h = 2e-5, ds <= 20) and
    # the rounding error eps * |L| / h ~ 1e-10 absolute.
    for c, ch in pick:
        h = 1e-3 if ch < 3 else 2e-5
        fp, fm = f0.clone(), f0.clone()
        fp[c, ch] += h
        fm[c, ch] -= h
        fd = (loss(fp).item() - loss(fm).item()) / (2 * h)
        tol = FD_RTOL * max(abs(fd), abs(grad[c, ch])) + 1e-15 * abs(L.item()) / h
        assert abs(fd - grad[c, ch]) <= tol, (c, ch, fd, grad[c, ch])


def test_adjoint_is_deterministic_and_checks_mode():
    _say("adjoint determinism and mode check")
    pos, rho = synthetic.uniform_mesh(5000, seed=14)
    pc = vt.ProjectionCloud(pos, rho, boundbox=[0, 100, 0, 100, 0, 100], periodic=True)
    s, e = _rays(24)
    dev = torch.device("cuda")
    ds_, de_ = torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev)
    g = torch.from_numpy(np.random.default_rng(5).standard_normal((len(s), 1))).to(dev)
    p = vt.Projection(ds_, de_)
    a = p.adjoint_tensor(pc._cloud, g, vt.ReductionMode.Sum).cpu().numpy()
    b = p.adjoint_tensor(pc._cloud, g, vt.ReductionMode.Sum).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    assert np.count_nonzero(a) > 0
    with pytest.raises(ValueError, match="adjoint"):
        p.adjoint_tensor(pc._cloud, g, vt.ReductionMode.Max)
