"""GPU parity: the HIP path against the oracle (oracle/vt_oracle.c, itself pinned
to the reference in test_oracle_pins.py) and, where present, the reference's
own build (oracle/_ref).

Bar: cell-id sequences, s_enter/s_exit bits and sum/max/min values
bit-identical; volume values within 1e-12 relative (GPU exp vs glibc exp);
kNN ids and r^2 bit-identical.
"""
import sys

import numpy as np
import pytest
import torch

from conftest import LATTICE_RAYS, cubic_lattice, grid_cloud
from oracle import port, ref

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import grid, synthetic  # noqa: E402

VOL_RTOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cmp_segments(pc, oc, s, e, mode="sum", vol_rtol=None):
    """GPU traced segments vs the oracle's, per ray, bit for bit."""
    _say(f"  compare {len(s)} rays mode={mode}")
    v, ids, se, sx, off = pc._cloud and vt.Projection(s, e).segments_tensor(
        pc._cloud, pc._validate_reduction(mode))
    v, ids, se, sx, off = (t.cpu().numpy() for t in (v, ids, se, sx, off))
    ov, oids, ose, osx, ooff = oc.detailed(s, e, mode)
    np.testing.assert_array_equal(off, ooff, err_msg="segment counts")
    np.testing.assert_array_equal(ids, oids, err_msg="cell ids")
    np.testing.assert_array_equal(se, ose, err_msg="s_enter")
    np.testing.assert_array_equal(sx, osx, err_msg="s_exit")
    if vol_rtol is None:
        np.testing.assert_array_equal(v, ov, err_msg="values")
    else:
        np.testing.assert_allclose(v, ov, rtol=vol_rtol, atol=0, err_msg="values")
    # the fused single-pass kernel gives the same values
    pv = pc.projection(s, e, reduction=mode)
    pv = pv.reshape(ov.shape)
    if vol_rtol is None:
        np.testing.assert_array_equal(pv, ov, err_msg="fused values")
    else:
        np.testing.assert_allclose(pv, ov, rtol=vol_rtol, atol=0)
    return off


def test_runtime_is_torchs():
    """One HIP runtime in the process: the library bound to torch's."""
    torch.zeros(1, device="cuda")
    with open("/proc/self/maps") as f:
        libs = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(libs) == 1, libs


@pytest.mark.parametrize("periodic", [False, True])
def test_knn_parity(periodic):
    rng = np.random.default_rng(11)
    pos = rng.random((20000, 3)) * 100
    rho = np.exp(rng.standard_normal(20000))
    bb = [0, 100, 0, 100, 0, 100]
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    pc = vt.PointCloud()
    pc.loadPoints(pos, rho, bb, periodic=periodic)
    pc.buildTree()
    q = rng.random((4000, 3)) * 100
    q[:1000] *= 0.02
    q[1000:1100] = pos[:100]  # exact hits
    for k in (1, 2, 3, 5, 8, 100):
        _say(f"  knn k={k} periodic={periodic}")
        ids, r2 = pc.query_tensor(torch.from_numpy(q), k)
        oi, orr = oc.query(q, k)
        np.testing.assert_array_equal(ids.cpu().numpy(), oi, err_msg=f"ids k={k}")
        np.testing.assert_array_equal(r2.cpu().numpy(), orr, err_msg=f"r2 k={k}")


def test_lattice_degenerate():
    """test_degenerate.py / test_traced_projection.py rays on the 5^3 lattice:
    vertices, faces, edges, parallel bisectors, case 1a, K_EXPAND."""
    pos, f, vol, bb = cubic_lattice()
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    oc = port.OracleCloud(pos, f, bb, vol=vol)
    s = np.array([r[0] for r in LATTICE_RAYS], float)
    e = np.array([r[1] for r in LATTICE_RAYS], float)
    off = _cmp_segments(pc, oc, s, e)
    assert (np.diff(off) > 0).all()
    vals = pc.projection(s, e)
    np.testing.assert_allclose(vals, np.linalg.norm(e - s, axis=1), rtol=1e-6)


def test_lattice_float32_positions():
    n = 10
    pos, f, vol, bb = cubic_lattice(n)
    off = np.array([500.0, 500.0, 500.1])
    p32 = (pos + off).astype(np.float32)
    p32 -= off.astype(np.float32)
    p = p32.astype(np.float64)
    pc = vt.ProjectionCloud(p, f, boundbox=bb, vol=vol)
    oc = port.OracleCloud(p, f, bb, vol=vol)
    s = np.array([[-0.1] * 3, [0.3, -0.1, 0.2]])
    e = np.array([[n - 0.9] * 3, [8.7, 9.2, 8.9]])
    _cmp_segments(pc, oc, s, e)


def test_galaxy_golden(galaxy, galaxy_golden):
    """test_vortrace.py:12-29 at decimal=14, and bit-equal to the oracle."""
    box = galaxy["box_size"]
    pc = vt.ProjectionCloud(galaxy["pos"], galaxy["dens"],
                            boundbox=[0., box, 0., box, 0., box])
    ext = [box / 2. - 75. / 2., box / 2. + 75. / 2.]
    img = pc.grid_projection(ext, 128, [0., box], None)
    assert img.shape == (128, 128)
    np.testing.assert_array_almost_equal(img, galaxy_golden, decimal=14)
    oc = port.OracleCloud(galaxy["pos"].astype(np.float64), galaxy["dens"],
                          [0., box, 0., box, 0., box])
    s, e = grid.generate_projection_grid(ext, 128, [0., box], None)
    np.testing.assert_array_equal(img.ravel(), oc.project(s.reshape(-1, 3),
                                                          e.reshape(-1, 3))[:, 0])
    if ref.available():
        rc = ref.RefCloud(galaxy["pos"].astype(np.float64), galaxy["dens"],
                          [0., box, 0., box, 0., box])
        np.testing.assert_array_equal(img.ravel(), rc.project(s.reshape(-1, 3),
                                                              e.reshape(-1, 3))[:, 0])
    s1 = np.array([0, 0, 0.])
    val, ids, sv, dv = pc.traced_projection(s1, np.array([box, box, box]))
    assert np.isscalar(val) or np.ndim(val) == 0


@pytest.mark.parametrize("periodic", [False, True])
def test_uniform_mesh(periodic):
    """C1-like uniform mesh (seed 1), all rays of a 48^2 grid."""
    pos, rho = synthetic.uniform_mesh(100_000, seed=1)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    s, e = grid.generate_projection_grid([0, 100], 48, [0, 100], None)
    _cmp_segments(pc, oc, s.reshape(-1, 3), e.reshape(-1, 3))


def test_clustered_periodic():
    pos, rho = synthetic.clustered_mesh(150_000, seed=3, workers=8)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)
    oc = port.OracleCloud(pos, rho, bb, periodic=True)
    s, e = grid.generate_projection_grid([0, 100], 48, [0, 100], None)
    _cmp_segments(pc, oc, s.reshape(-1, 3), e.reshape(-1, 3))
    # rotated cameras: arbitrary directions through the periodic box
    for (vs, ve) in synthetic.camera_rays(4, 3, sizes=(16,)):
        _cmp_segments(pc, oc, vs, ve)


def test_periodic_grid_rays():
    """test_periodic.py:126-196."""
    pos, f, bb = grid_cloud(5.0, 0.5, 2.5)
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, periodic=True)
    oc = port.OracleCloud(pos, f, bb, periodic=True)
    s = np.array([[0.01, 2.5, 2.5], [0.5, 2.5, 2.5], [0.0, 0.0, 0.0], [0, 1.3, 0]])
    e = np.array([[4.99, 2.5, 2.5], [4.5, 2.5, 2.5], [5.0, 5.0, 5.0], [5, 1.3, 5]])
    _cmp_segments(pc, oc, s, e)
    dens, ids, _, _ = pc.traced_projection(s[0], e[0])
    assert len(ids) == len(set(ids))
    np.testing.assert_allclose(dens, 2.5 * np.linalg.norm(e[0] - s[0]), rtol=1e-10)


def test_volume_render():
    rng = np.random.default_rng(123)
    n = 5000
    pos = rng.random((n, 3)) * 100
    f4 = np.column_stack([rng.random(n), rng.random(n), rng.random(n),
                          0.01 + 0.05 * rng.random(n)])
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, f4, boundbox=bb)
    oc = port.OracleCloud(pos, f4, bb)
    s, e = grid.generate_projection_grid([5, 95], 24, [0, 100], None)
    _cmp_segments(pc, oc, s.reshape(-1, 3), e.reshape(-1, 3), "volume", VOL_RTOL)
    img = pc.grid_projection([5, 95], 24, [0, 100], None, reduction="volume")
    assert img.shape == (24, 24, 3) and (img >= 0).all()
    # uniform medium: C (1 - exp(-alpha L))  (test_volume_render.py:30-47)
    pcu = vt.ProjectionCloud(pos, np.tile([0.8, 0.4, 0.2, 0.01], (n, 1)), boundbox=bb)
    rgb, _, _, ds = pcu.traced_projection(np.array([50, 50, 0.1]),
                                          np.array([50, 50, 99.9]), reduction="volume")
    np.testing.assert_allclose(rgb, np.array([0.8, 0.4, 0.2]) * (1 - np.exp(-0.01 * ds.sum())),
                               rtol=1e-10)


def test_multifield_and_extrema(galaxy):
    box = galaxy["box_size"]
    d = galaxy["dens"].astype(np.float64)
    f = np.column_stack([d, 2 * d, 3 * d])
    bb = [0., box, 0., box, 0., box]
    pc = vt.ProjectionCloud(galaxy["pos"], f, boundbox=bb)
    oc = port.OracleCloud(galaxy["pos"].astype(np.float64), f, bb)
    s, e = grid.generate_projection_grid([box / 2 - 30, box / 2 + 30], 32, [0, box], None)
    s, e = s.reshape(-1, 3), e.reshape(-1, 3)
    for mode in ("sum", "max", "min"):
        _cmp_segments(pc, oc, s, e, mode)
    v = pc.grid_projection([box / 2 - 30, box / 2 + 30], 32, [0, box], None)
    assert v.shape == (32, 32, 3)
    np.testing.assert_array_equal(v[..., 1], 2 * v[..., 0])
    f5 = np.column_stack([d, d, d, d, d])  # generic field count path
    pc5 = vt.ProjectionCloud(galaxy["pos"], f5, boundbox=bb)
    oc5 = port.OracleCloud(galaxy["pos"].astype(np.float64), f5, bb)
    _cmp_segments(pc5, oc5, s[:200], e[:200])


def test_batch_equals_single_and_formats():
    pos, f, vol, bb = cubic_lattice()
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    s = np.array([r[0] for r in LATTICE_RAYS[:6]], float)
    e = np.array([r[1] for r in LATTICE_RAYS[:6]], float)
    vals, cl, sl, dl = pc.traced_projection(s, e)
    for i in range(len(s)):
        v1, c1, s1, d1 = pc.traced_projection(s[i], e[i])
        np.testing.assert_array_equal(cl[i], c1)
        np.testing.assert_array_equal(sl[i], s1)
        np.testing.assert_array_equal(dl[i], d1)
        np.testing.assert_array_equal(vals[i], v1)
    out = pc.traced_projection(s, e, flatten=True)
    assert len(out) == 5 and out[4][0] == 0 and out[4][-1] == len(out[1])
    _, _, smid, _ = pc.traced_projection(s[0], e[0], return_midpoint=True)
    _, _, sent, _ = pc.traced_projection(s[0], e[0], return_midpoint=False)
    assert not np.array_equal(smid, sent)


def test_empty_and_degenerate_inputs():
    pos, f, vol, bb = cubic_lattice()
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    out = pc.projection(np.zeros((0, 3)), np.zeros((0, 3)))
    assert out.shape == (0,)
    # zero-length ray: one segment of length 0 (non-periodic)
    v, ids, _, ds = pc.traced_projection(np.array([1.2, 1.3, 1.4]), np.array([1.2, 1.3, 1.4]))
    assert len(ids) == 1 and ds[0] == 0.0
    with pytest.raises(ValueError, match="non-finite"):
        pc.projection(np.array([[np.nan, 0, 0]]), np.array([[1.0, 1, 1]]))
    # one-point cloud (test_cpp.cpp:110-134)
    one = vt.ProjectionCloud(np.array([[0.5, 0.5, 0.5]]), np.array([1.0]),
                             boundbox=[0, 1, 0, 1, 0, 1])
    np.testing.assert_allclose(one.projection(np.array([[0.5, 0.5, 0.0]]),
                                              np.array([[0.5, 0.5, 1.0]])), [1.0])


def test_knn_tiny_periodic_duplicates():
    """A 12-point periodic cloud in a unit box: images return the same
    particle several times and must be merged (src/pointcloud.cpp:340-353)."""
    rng = np.random.default_rng(5)
    pos = rng.random((12, 3))
    bb = [0, 1, 0, 1, 0, 1]
    pc = vt.PointCloud()
    pc.loadPoints(pos, np.ones(12), bb, periodic=True)
    pc.buildTree()
    oc = port.OracleCloud(pos, np.ones(12), bb, periodic=True)
    q = rng.random((200, 3))
    for k in (1, 2, 3, 10):
        ids, r2 = pc.query_tensor(torch.from_numpy(q), k)
        oi, orr = oc.query(q, k)
        np.testing.assert_array_equal(ids.cpu().numpy(), oi)
        np.testing.assert_array_equal(r2.cpu().numpy(), orr)


@pytest.mark.parametrize("scale", [1e-3, 1e5, 1e7])
@pytest.mark.parametrize("periodic", [False, True])
def test_bounded_searches_across_scales(scale, periodic):
    """The split-point classification searches are bounded by the nearer
    cell's r^2 plus 2 kTol (kTol = 1e-9, absolute): at box 1e-3 the bound is
    mostly the 2 kTol; at 1e5 2 kTol is near the resolution of r^2; at 1e7 it
    is below it and the searches run unbounded.  Bits must match everywhere."""
    _say(f"bounded searches scale={scale} periodic={periodic}")
    pos, rho = synthetic.uniform_mesh(3000, seed=15)
    pos = pos * (scale / 100.0)
    bb = [0, scale, 0, scale, 0, scale]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    s, e = grid.generate_projection_grid([0.0, scale], 12, [0.0, scale], None)
    off = _cmp_segments(pc, oc, s.reshape(-1, 3), e.reshape(-1, 3))
    assert off[-1] > 2000


@pytest.mark.parametrize("periodic", [False, True])
def test_zero_length_ray_among_others(periodic):
    """A zero-length ray (start == end) in a call with many ordinary rays: one
    segment of length 0 in its start cell (src/ray.cpp:212-218) without
    periodic images; with them the reference's dyadic probe of a zero-length
    ray divides by zero, so only the non-periodic case is compared."""
    pos, rho = synthetic.clustered_mesh(20_000, seed=93, n_clumps=6)
    bb = [0, 100, 0, 100, 0, 100]
    s, e = grid.generate_projection_grid([0, 100], 40, [0, 100], None)
    s, e = s.reshape(-1, 3).copy(), e.reshape(-1, 3).copy()
    e[9] = s[9]
    e[700] = s[700] = (37.5, 61.25, 48.0)
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    if periodic:
        keep = np.ones(len(s), bool)
        keep[[9, 700]] = False
        s, e = s[keep], e[keep]
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    _cmp_segments(pc, oc, s, e)
