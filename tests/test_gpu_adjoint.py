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

from conftest import LATTICE_RAYS, cubic_lattice
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


def _recorded_case(case):
    dev = torch.device("cuda")
    rng = np.random.default_rng(31)
    if case == "lattice":
        # The reference's degenerate rays, repeated so that many lanes and
        # record chunks carry them; K_EXPAND rays run in the second launch.
        pos, _, vol, bb = cubic_lattice()
        f = 0.5 + rng.random(len(pos))
        pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
        s = np.array([r[0] for r in LATTICE_RAYS] * 40, float)
        e = np.array([r[1] for r in LATTICE_RAYS] * 40, float)
        mode = vt.ReductionMode.Sum
    else:
        pos, rho = synthetic.clustered_mesh(40_000, seed=32, n_clumps=12)
        bb = [0, 100, 0, 100, 0, 100]
        if case == "volume":
            f = np.empty((len(pos), 4))
            f[:, :3] = 0.1 + 0.8 * rng.random((len(pos), 3))
            f[:, 3] = 0.02 + 0.05 * rng.random(len(pos))
            mode = vt.ReductionMode.VolumeRender
        else:
            f = np.column_stack([rho, rng.random(len(pos))])
            mode = vt.ReductionMode.Sum
        pc = vt.ProjectionCloud(pos, f, boundbox=bb, periodic=True)
        s, e = _rays(40)
    return pc, f, torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev), mode


@pytest.mark.parametrize("case", ["sum", "volume", "lattice"])
def test_recorded_adjoint_matches_segment_walk(case):
    """One-walk forward/backward (vt_project_recorded, vt_adjoint_recorded)
    against the projection and the segment-walk adjoint: the same bits;
    an overflowing pool reports itself and leaves image and counts whole."""
    _say(f"recorded adjoint case={case}")
    pc, _, ds_, de_, mode = _recorded_case(case)
    dev = ds_.device
    p = vt.Projection(ds_, de_)
    n = p.ngrid
    c_ref = torch.empty(n, dtype=torch.int64, device=dev)
    img = p.project_tensor(pc._cloud, mode, counts=c_ref)
    total = int(c_ref.sum().item())
    counts = torch.empty(n, dtype=torch.int64, device=dev)
    pool = torch.empty((total + 2 * 131072 * 32, vt.core.SEGREC_BYTES), dtype=torch.uint8,
                       device=dev)
    # The pool is not cleared by the call (each lane marks the unfilled rest
    # of its last chunk): fill it with stale records that pass the reorder's
    # checks (ray j mod n, place 0, cell 0, ds 1e300) so that one left
    # unmarked would change the gradient.
    stale = pool.view(torch.int32)
    stale[:, 0] = 0
    stale[:, 1] = (torch.arange(pool.shape[0], device=dev) % n).to(torch.int32)
    stale[:, 2] = 0
    pool.view(torch.float64)[:, 2] = 1e300
    img_r, used = p.project_recorded_tensor(pc._cloud, mode, counts, pool)
    deferred = vt._lib.last_call_stats().deferred_rays
    _say(f"  {n} rays, {total} segments, {used} records, {deferred} deferred rays")
    if case == "lattice":
        assert deferred > 0
    assert total <= used <= pool.shape[0]
    assert torch.equal(img_r, img) and torch.equal(counts, c_ref)
    g = torch.from_numpy(np.random.default_rng(33).standard_normal(img.shape)).to(dev)
    a = p.adjoint_tensor(pc._cloud, g, mode, counts=c_ref)
    b = p.adjoint_recorded_tensor(pc._cloud, g, mode, counts, pool, used)
    assert torch.equal(a, b)
    assert int(torch.count_nonzero(b)) > 0
    small = torch.empty((total // 2, vt.core.SEGREC_BYTES), dtype=torch.uint8, device=dev)
    counts.fill_(-1)
    img_o, used_o = p.project_recorded_tensor(pc._cloud, mode, counts, small)
    assert used_o == -1
    assert torch.equal(img_o, img) and torch.equal(counts, c_ref)
    with pytest.raises(ValueError, match="overflowed"):
        p.adjoint_recorded_tensor(pc._cloud, g, mode, counts, small, used_o)


def test_autograd_records_after_first_pass():
    """The autograd wrapper records segments from its second forward pass on
    (pool sized from the first); gradients keep their bits."""
    _say("autograd recorded passes")
    pc, f_np, ds_, de_, _ = _recorded_case("sum")
    f0 = torch.from_numpy(f_np).to(ds_.device)
    w = torch.from_numpy(np.random.default_rng(35).standard_normal((ds_.shape[0], 2))).to(
        ds_.device)
    grads = []
    for _ in range(3):
        f = f0.clone().requires_grad_(True)
        (vt.autograd.project(pc, f, ds_, de_, "sum") * w).sum().backward()
        grads.append(f.grad)
        states = list(vt.autograd._ray_states(pc).values())
        assert len(states) == 1 and states[0].estimate > 0
        assert pc._adjoint_path == ("walk" if len(grads) == 1 else "recorded")
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[1], grads[2])
    assert int(torch.count_nonzero(grads[0])) > 0


def test_autograd_first_pass_sampled_pool(monkeypatch):
    """The first pass over a ray set sizes its record pool from a sample of
    the rays and records; a pool that falls short leaves the backward pass to
    walk the rays.  The gradient has the same bits either way."""
    _say("autograd first pass, sampled pool")
    pc, f_np, ds_, de_, _ = _recorded_case("sum")
    monkeypatch.setattr(vt.autograd, "_SAMPLE_RAYS", 256)  # every 6th of 1600 rays
    f0 = torch.from_numpy(f_np).to(ds_.device)
    w = torch.from_numpy(np.random.default_rng(36).standard_normal((ds_.shape[0], 2))).to(
        ds_.device)

    def first_pass():
        vt.autograd._ray_states(pc).clear()
        f = f0.clone().requires_grad_(True)
        (vt.autograd.project(pc, f, ds_, de_, "sum") * w).sum().backward()
        return f.grad, pc._adjoint_path

    g1, path1 = first_pass()
    assert path1 == "recorded"
    # a pool far too small for the sampled estimate: overflow, then a walk
    monkeypatch.setattr(vt.autograd, "_pool_capacity", lambda est, lanes: max(1, est // 100))
    g2, path2 = first_pass()
    assert path2 == "walk"
    assert torch.equal(g1, g2) and int(torch.count_nonzero(g1)) > 0
