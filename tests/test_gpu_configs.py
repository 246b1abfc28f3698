"""GPU parity at the benchmark workloads, with the benchmark's own meshes,
seeds and rays (bench.py: make_workload, volume_fields, fit_target).

* uniform 100k (seed 1), 128^2 z-rays, sum: every ray's segments (cell ids,
  s_enter/s_exit bits) and value equal the oracle's.
* clustered 2M periodic (seed 2), 512^2 z-rays, sum: the whole image
  bit-identical to the oracle's; segments of every 16th ray.
* the same mesh with RGBA fields, 1024^2 z-rays, volume: every 4th ray's
  value within 1e-12 relative; segments of every 256th ray.
* clustered 4M periodic (seed 4), 64 cameras in ONE batched launch (rays
  built on the device): a strided sample of every view bit-identical to the
  oracle; segments of a sparser sample.
* clustered 8M periodic (seed 5), 16 views at 512^2, fit step: forward values
  on a strided sample, the gradient of a loss weighting 8 rays equal to the
  Jacobian rows built from the oracle's segment lists (bit for bit), and
  central differences of the fit loss on 64 cells within 1e-6 relative.
* clustered 1M periodic (seed 6), RGBA, 256^2, volume fit: the gradient of
  a loss weighting 8 rays against the closed-form derivative of
  reduce_volume_render built from the oracle's segments (1e-12 relative to
  the magnitude of the terms), and central differences on 64 cells.

The oracle is oracle/vt_oracle.c (pinned to the reference in
test_oracle_pins.py).  Meshes are built once per module.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port

pytestmark = pytest.mark.gpu

import bench  # noqa: E402
import vortrace_amd as vt  # noqa: E402
from vortrace_amd import grid, synthetic  # noqa: E402

BOX = [0, 100, 0, 100, 0, 100]
VOL_RTOL = 1e-12
FD_RTOL = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cmp_segments(pc, oc, s, e, mode, vol_rtol=None, what=""):
    """The GPU's traced segments of rays (s, e) against the oracle's, per ray,
    bit for bit; values bit-identical (or within vol_rtol)."""
    _say(f"  {what}: segments of {len(s)} rays ({mode})")
    v, ids, se, sx, off = vt.Projection(s, e).segments_tensor(
        pc._cloud, pc._validate_reduction(mode))
    v, ids, se, sx, off = (t.cpu().numpy() for t in (v, ids, se, sx, off))
    ov, oids, ose, osx, ooff = oc.detailed(s, e, mode)
    np.testing.assert_array_equal(off, ooff, err_msg=f"{what}: segment counts")
    np.testing.assert_array_equal(ids, oids, err_msg=f"{what}: cell ids")
    np.testing.assert_array_equal(se, ose, err_msg=f"{what}: s_enter")
    np.testing.assert_array_equal(sx, osx, err_msg=f"{what}: s_exit")
    if vol_rtol is None:
        np.testing.assert_array_equal(v, ov, err_msg=f"{what}: values")
    else:
        np.testing.assert_allclose(v, ov, rtol=vol_rtol, atol=0, err_msg=f"{what}: values")
    return int(off[-1])


def _cmp_values(got, oc, s, e, mode, vol_rtol=None, what=""):
    want = oc.project(s, e, mode)
    got = np.asarray(got).reshape(want.shape)
    if vol_rtol is None:
        np.testing.assert_array_equal(got, want, err_msg=f"{what}: values")
    else:
        np.testing.assert_allclose(got, want, rtol=vol_rtol, atol=0,
                                   err_msg=f"{what}: values")


# --------------------------------------------------------------- uniform 100k

def test_uniform100k_128():
    pos, rho, views, _ = bench.make_workload("uniform100k")
    s, e = views[0]
    pc = vt.ProjectionCloud(pos, rho, boundbox=BOX, periodic=False)
    oc = port.OracleCloud(pos, rho, BOX, periodic=False)
    img = pc.grid_projection([0.0, 100.0], 128, [0.0, 100.0], None)
    _cmp_values(img.ravel(), oc, s, e, "sum", what="uniform100k image")
    n = _cmp_segments(pc, oc, s, e, "sum", what="uniform100k")
    _say(f"  uniform100k: {len(s)} rays, {n} crossings bit-identical")
    assert len(s) == 128 * 128 and n > 1_000_000


# ---------------------------------------------------------------- clustered 2M

@pytest.fixture(scope="module")
def clustered2m():
    _say("building the 2M clustered mesh (seed 2)")
    pos, rho, views, _ = bench.make_workload("clustered2m")
    return pos, rho, views[0]


def test_clustered2m_512_sum(clustered2m):
    pos, rho, (s, e) = clustered2m
    pc = vt.ProjectionCloud(pos, rho, boundbox=BOX, periodic=True)
    oc = port.OracleCloud(pos, rho, BOX, periodic=True)
    img = pc.grid_projection([0.0, 100.0], 512, [0.0, 100.0], None)
    assert img.shape == (512, 512)
    _say("  clustered2m: whole 512^2 image against the oracle")
    _cmp_values(img.ravel(), oc, s, e, "sum", what="clustered2m image")
    # the device-resident call the benchmark times gives the same bits
    dev = torch.device("cuda")
    pc.set_ray_rows(512)
    out = pc.projection_tensor(torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev))
    pc.set_ray_rows(0)
    np.testing.assert_array_equal(out.cpu().numpy()[:, 0], img.ravel())
    n = _cmp_segments(pc, oc, s[::16], e[::16], "sum", what="clustered2m every 16th ray")
    assert n > 1_000_000


def test_clustered2m_1024_volume(clustered2m):
    pos, rho, _ = clustered2m
    rgba, kappa = bench.volume_fields(pos, rho)
    _say(f"  clustered2m volume: kappa {kappa:.6e}")
    s, e = grid.generate_projection_grid([0.0, 100.0], 1024, [0.0, 100.0], None)
    s, e = s.reshape(-1, 3), e.reshape(-1, 3)
    pc = vt.ProjectionCloud(pos, rgba, boundbox=BOX, periodic=True)
    oc = port.OracleCloud(pos, rgba, BOX, periodic=True)
    img = pc.grid_projection([0.0, 100.0], 1024, [0.0, 100.0], None, reduction="volume")
    assert img.shape == (1024, 1024, 3)
    flat = img.reshape(-1, 3)
    _say("  clustered2m volume: every 4th ray against the oracle")
    _cmp_values(flat[::4], oc, s[::4], e[::4], "volume", VOL_RTOL, "clustered2m volume")
    _cmp_segments(pc, oc, s[::256], e[::256], "volume", VOL_RTOL,
                  "clustered2m volume every 256th ray")
    # colours come from narrow bumps in log10 rho: many rays stay near black,
    # the image as a whole does not
    assert flat.max() > 0.01 and (flat >= 0.0).all()


# ---------------------------------------------------------------- multiview 4M

def test_multiview4m_one_launch():
    pos, rho, views, params = bench.make_workload("multiview4m")
    dev = torch.device("cuda")
    pc = vt.ProjectionCloud(pos, rho, boundbox=BOX, periodic=True)
    d_start, d_end, sizes = grid.generate_projection_grids_device(params, dev)
    assert sizes == [len(v[0]) for v in views]
    d_views = list(zip(torch.split(d_start, sizes), torch.split(d_end, sizes)))
    _say(f"  multiview4m: {len(sizes)} views, {sum(sizes)} rays in one launch")
    out = torch.cat(pc.batch_projection_tensor(d_views, reduction="sum"))
    st = vt._lib.last_call_stats()
    _say(f"  multiview4m: {st.segments} crossings, {st.kernel_ms:.1f} ms")
    assert st.segments > 1_000_000_000
    got = out.cpu().numpy()[:, 0]
    hs, he = d_start.cpu().numpy(), d_end.cpu().numpy()
    same = all(np.array_equal(hs[o:o + n], v[0]) and np.array_equal(he[o:o + n], v[1])
               for v, o, n in zip(views, np.cumsum([0] + sizes[:-1]), sizes))
    _say(f"  multiview4m: device rays identical to numpy's: {same}")
    oc = port.OracleCloud(pos, rho, BOX, periodic=True)
    # every 97th ray of the concatenation: a sample of every view
    idx = np.arange(0, len(hs), 97)
    assert len(np.unique(np.searchsorted(np.cumsum(sizes), idx, side="right"))) == len(sizes)
    _cmp_values(got[idx], oc, hs[idx], he[idx], "sum", what="multiview4m sample")
    _cmp_segments(pc, oc, hs[::1501], he[::1501], "sum", what="multiview4m every 1501st ray")


# --------------------------------------------------------------------- fit 8M

def _weighted_rays(n, rays, rng, nout=1):
    w = np.zeros((n, nout))
    w[rays] = 0.5 + rng.random((len(rays), nout))
    return w


def _sum_rows(oc, s, e, rays, w, npart):
    """Gradient of sum_r w_r proj_r from the oracle's segments of `rays`: every
    cell adds w_r * ds over (ray, segment) in ray-major order, from 0.0."""
    _, ids, se, sx, off = oc.detailed(s[rays], e[rays], "sum")
    g = np.zeros(npart)
    for k, r in enumerate(rays):
        for j in range(off[k], off[k + 1]):
            g[ids[j]] = g[ids[j]] + w[r, 0] * (sx[j] - se[j])
    return g, int(off[-1])


def test_fit8m_forward_adjoint():
    pos, rho, views, params = bench.make_workload("fit8m")
    dev = torch.device("cuda")
    pc = vt.ProjectionCloud(pos, rho, boundbox=BOX, periodic=True)
    d_start, d_end, sizes = grid.generate_projection_grids_device(params, dev)
    hs, he = d_start.cpu().numpy(), d_end.cpu().numpy()
    n = len(hs)
    _say(f"  fit8m: {len(sizes)} views, {n} rays")
    oc = port.OracleCloud(pos, rho, BOX, periodic=True)
    f0 = torch.from_numpy(rho).to(dev)
    # forward of the fit (autograd path) against the oracle on a sample
    p = vt.autograd.project(pc, f0, d_start, d_end, "sum")
    idx = np.arange(0, n, 61)
    _cmp_values(p.cpu().numpy()[idx], oc, hs[idx], he[idx], "sum", what="fit8m forward")
    # Jacobian rows: a loss weighting 8 rays, through autograd; every pass
    # records its segments (one walk for forward and backward)
    rng = np.random.default_rng(57)
    rays = np.sort(rng.choice(n, 8, replace=False))
    w = _weighted_rays(n, rays, rng)
    wt = torch.from_numpy(w[:, 0]).to(dev)
    want, nseg = _sum_rows(oc, hs, he, rays, w, len(rho))
    for k in range(2):
        f = f0.clone().requires_grad_(True)
        (vt.autograd.project(pc, f, d_start, d_end, "sum") * wt).sum().backward()
        _say(f"  fit8m: Jacobian rows of rays {rays.tolist()} ({nseg} segments), "
             f"pass {k + 1} ({pc._adjoint_path})")
        np.testing.assert_array_equal(f.grad.cpu().numpy(), want,
                                      err_msg=f"fit8m Jacobian rows, pass {k + 1}")
        # the first pass records too: its pool is sized from a sample of the rays
        assert pc._adjoint_path == "recorded"
    # the fit's loss and central differences on 64 cells
    target = bench.fit_target(pc, rho, d_start, d_end, "sum")
    grads = []
    for _ in range(2):
        f = f0.clone().requires_grad_(True)
        ((vt.autograd.project(pc, f, d_start, d_end, "sum") - target) ** 2).sum().backward()
        grads.append(f.grad)
    assert torch.equal(grads[0], grads[1])
    fd = bench.fd_check(pc, rho, grads[1], target, d_start, d_end, "sum")
    _say(f"  fit8m: central differences {fd}")
    assert fd["cells"] == 64 and fd["all_within_1e-6"], fd


# ---------------------------------------------------------------- fit 1M volume

def _volume_rows(oc, s, e, rays, w, rgba):
    """Gradient of sum_r w_r . proj_r (volume) from the oracle's segments:
    with op_i = 1 - exp(-a_i ds_i), T_i = prod_{j<i} (1 - op_j),
      d C_ch / d col_i,ch = T_i op_i
      d C_ch / d a_i      = T_i ds_i exp(-a_i ds_i) col_i,ch
                            - ds_i sum_{j>i} T_j op_j col_j,ch
    (the second term: every later T_j carries the factor 1 - op_i).  Also
    returns the magnitude of the terms each entry sums, for the tolerance."""
    _, ids, se, sx, off = oc.detailed(s[rays], e[rays], "volume")
    npart = len(rgba)
    g = np.zeros((npart, 4))
    mag = np.zeros((npart, 4))
    for k, r in enumerate(rays):
        c = ids[off[k]:off[k + 1]]
        ds = sx[off[k]:off[k + 1]] - se[off[k]:off[k + 1]]
        col, a = rgba[c, :3], rgba[c, 3]
        ex = np.exp(-a * ds)
        op = 1.0 - ex
        T = np.concatenate([[1.0], np.cumprod(ex)[:-1]])
        fwd = (T * op)[:, None] * col                       # per segment, per channel
        later = np.cumsum(fwd[::-1], axis=0)[::-1] - fwd     # sum over j > i
        later_abs = np.cumsum(np.abs(fwd)[::-1], axis=0)[::-1] - np.abs(fwd)
        dcol = (T * op)[:, None] * w[r][None, :]
        da_terms = (T * ds * ex)[:, None] * col - ds[:, None] * later
        da = (da_terms * w[r][None, :]).sum(axis=1)
        da_mag = ((T * ds * ex)[:, None] * col + ds[:, None] * later_abs) @ np.abs(w[r])
        np.add.at(g, (c, slice(0, 3)), dcol)
        np.add.at(g[:, 3], c, da)
        np.add.at(mag, (c, slice(0, 3)), np.abs(dcol))
        np.add.at(mag[:, 3], c, da_mag)
    return g, mag


def test_fit1m_volume_forward_adjoint():
    pos, rho, views, _ = bench.make_workload("fit1m_volume")
    rgba, kappa = bench.volume_fields(pos, rho)
    s, e = views[0]
    dev = torch.device("cuda")
    d_start, d_end = torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev)
    pc = vt.ProjectionCloud(pos, rgba, boundbox=BOX, periodic=True)
    oc = port.OracleCloud(pos, rgba, BOX, periodic=True)
    f0 = torch.from_numpy(rgba).to(dev)
    p = vt.autograd.project(pc, f0, d_start, d_end, "volume")
    _say(f"  fit1m_volume: kappa {kappa:.6e}, {len(s)} rays")
    _cmp_values(p.cpu().numpy(), oc, s, e, "volume", VOL_RTOL, "fit1m_volume forward")
    rng = np.random.default_rng(58)
    rays = np.sort(rng.choice(len(s), 8, replace=False))
    w = _weighted_rays(len(s), rays, rng, 3)
    wt = torch.from_numpy(w).to(dev)
    want, mag = _volume_rows(oc, s, e, rays, w, rgba)
    for k in range(2):
        f = f0.clone().requires_grad_(True)
        (vt.autograd.project(pc, f, d_start, d_end, "volume") * wt).sum().backward()
        got = f.grad.cpu().numpy()
        err = np.abs(got - want)
        bad = err > VOL_RTOL * mag
        _say(f"  fit1m_volume: Jacobian rows pass {k + 1}: max err/mag "
             f"{float(np.max(err / np.maximum(mag, 1e-300))):.2e}")
        assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], want[bad][:5])
        # the first pass records too, into a pool sized from a sample of the rays
        assert pc._adjoint_path == "recorded"
    target = bench.fit_target(pc, rgba, d_start, d_end, "volume")
    f = f0.clone().requires_grad_(True)
    ((vt.autograd.project(pc, f, d_start, d_end, "volume") - target) ** 2).sum().backward()
    fd = bench.fd_check(pc, rgba, f.grad, target, d_start, d_end, "volume")
    _say(f"  fit1m_volume: central differences {fd}")
    assert fd["cells"] == 64 and fd["all_within_1e-6"], fd
