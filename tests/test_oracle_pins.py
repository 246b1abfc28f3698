"""Pins of the oracle (oracle/vt_oracle.c) -- CPU only.

The oracle is what the GPU path is checked against, so it is itself checked:
  * against the reference's own C++ core (oracle/_ref, built from the
    reference sources) bit for bit: kd-tree, kNN ids and r^2, per-ray segment
    ids and s values, reduced values;
  * against the reference's golden image (decimal=14, test_vortrace.py:29);
  * against closed forms and brute force that a plausible mistake would break
    (ray-length integrals, volume C(1 - e^{-aL}), 27-image brute-force kNN,
    exact linearity in the fields, telescoping segments).
"""
import numpy as np
import pytest

from conftest import LATTICE_RAYS, cubic_lattice, grid_cloud
from oracle import port, ref
from vortrace_amd import grid

needs_ref = pytest.mark.skipif(not ref.available(),
                               reason="reference build oracle/_ref not present")


def _clouds(pos, fields, bb, vol=None, periodic=False):
    return (port.OracleCloud(pos, fields, bb, vol=vol, periodic=periodic),
            ref.RefCloud(pos, fields, bb, vol=vol, periodic=periodic))


def _same_detailed(a, b, s, e, mode="sum"):
    ra, rb = a.detailed(s, e, mode), b.detailed(s, e, mode)
    for x, y, name in zip(ra, rb, ["values", "cell_ids", "s_enter", "s_exit", "offsets"]):
        np.testing.assert_array_equal(x, y, err_msg=name)


def _same_tree(a, b):
    va, ba, na = a.tree()
    vb, bb, nb = ref.parse_tree_bytes(b.tree_bytes())
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(ba, bb)
    assert na == nb


def _galaxy_rays(box, npix=128):
    ext = [box / 2. - 37.5, box / 2. + 37.5]
    s, e = grid.generate_projection_grid(ext, npix, [0., box], None)
    return s.reshape(-1, 3), e.reshape(-1, 3)


def test_golden_image(galaxy, galaxy_golden):
    """test_vortrace.py:12-29: 128^2 projection of the galaxy at decimal=14."""
    box = galaxy["box_size"]
    c = port.OracleCloud(galaxy["pos"].astype(np.float64), galaxy["dens"],
                         [0., box, 0., box, 0., box])
    s, e = _galaxy_rays(box)
    img = c.project(s, e)[:, 0].reshape(128, 128)
    np.testing.assert_array_almost_equal(img, galaxy_golden, decimal=14)


@needs_ref
def test_galaxy_matches_reference(galaxy):
    box = galaxy["box_size"]
    a, b = _clouds(galaxy["pos"].astype(np.float64), galaxy["dens"],
                   [0., box, 0., box, 0., box])
    assert a.warnings == b.warnings and a.npart == b.npart
    np.testing.assert_array_equal(a.orig_ids, b.orig_ids)
    _same_tree(a, b)
    s, e = _galaxy_rays(box, 64)
    _same_detailed(a, b, s, e)


@needs_ref
def test_lattice_degenerate_matches_reference():
    pos, f, vol, bb = cubic_lattice()
    a, b = _clouds(pos, f, bb, vol=vol)
    _same_tree(a, b)
    s = np.array([r[0] for r in LATTICE_RAYS], float)
    e = np.array([r[1] for r in LATTICE_RAYS], float)
    _same_detailed(a, b, s, e)


@needs_ref
def test_lattice_float32_positions_matches_reference():
    """test_degenerate.py:94-116: float32-quantised lattice, diagonal ray."""
    n = 10
    pos, f, vol, bb = cubic_lattice(n)
    off = np.array([500.0, 500.0, 500.1])
    p32 = (pos + off).astype(np.float32)
    p32 -= off.astype(np.float32)
    a, b = _clouds(p32.astype(np.float64), f, bb, vol=vol)
    s = np.array([[-0.1] * 3, [0.3, -0.1, 0.2]])
    e = np.array([[n - 0.9] * 3, [8.7, 9.2, 8.9]])
    _same_detailed(a, b, s, e)
    v = a.project(s[:1], e[:1])[0, 0]
    np.testing.assert_allclose(v, np.linalg.norm(e[0] - s[0]), rtol=1e-3)


@needs_ref
@pytest.mark.parametrize("periodic", [False, True])
def test_uniform_random_matches_reference(periodic):
    rng = np.random.default_rng(11)
    pos = rng.random((20000, 3)) * 100
    rho = np.exp(rng.standard_normal(20000))
    a, b = _clouds(pos, rho, [0, 100, 0, 100, 0, 100], periodic=periodic)
    _same_tree(a, b)
    s, e = grid.generate_projection_grid([0, 100], 24, [0, 100], None)
    _same_detailed(a, b, s.reshape(-1, 3), e.reshape(-1, 3))
    q = rng.random((3000, 3)) * 100
    q[:1000] *= 0.02  # near the x=y=z=0 corner: many image searches
    for k in (1, 2, 3, 5, 8):
        ia, ra = a.query(q, k)
        ib, rb = b.query(q, k)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(ra, rb)


@needs_ref
def test_periodic_grid_rays_match_reference():
    """test_periodic.py:126-196: grid clouds, rays along / near faces."""
    pos, f, bb = grid_cloud(5.0, 0.5)
    a, b = _clouds(pos, f, bb, periodic=True)
    s = np.array([[0.01, 2.5, 2.5], [0.5, 2.5, 2.5], [0.0, 0.0, 0.0], [0, 1.3, 0]])
    e = np.array([[4.99, 2.5, 2.5], [4.5, 2.5, 2.5], [5.0, 5.0, 5.0], [5, 1.3, 5]])
    _same_detailed(a, b, s, e)


@needs_ref
def test_clustered_periodic_matches_reference():
    from vortrace_amd import synthetic
    pos, rho = synthetic.clustered_mesh(60000, seed=3, workers=4)
    a, b = _clouds(pos, rho, [0, 100, 0, 100, 0, 100], periodic=True)
    _same_tree(a, b)
    s, e = grid.generate_projection_grid([0, 100], 20, [0, 100], None)
    _same_detailed(a, b, s.reshape(-1, 3), e.reshape(-1, 3))


@needs_ref
def test_volume_and_extrema_match_reference():
    rng = np.random.default_rng(123)
    n = 3000
    pos = rng.random((n, 3)) * 100
    f4 = np.column_stack([rng.random(n), rng.random(n), rng.random(n),
                          0.01 + 0.05 * rng.random(n)])
    a, b = _clouds(pos, f4, [0, 100, 0, 100, 0, 100])
    s, e = grid.generate_projection_grid([5, 95], 12, [0, 100], None)
    for mode in ("volume", "max", "min", "sum"):
        _same_detailed(a, b, s.reshape(-1, 3), e.reshape(-1, 3), mode)


def test_periodic_knn_brute_force():
    """test_periodic.py:21-42,93-124: 27-image brute force, ids and r^2."""
    box = 10.0
    rng = np.random.RandomState(123)
    pos = rng.uniform(0.0, box, size=(200, 3))
    c = port.OracleCloud(pos, np.ones(200), [0, box, 0, box, 0, box], periodic=True)
    for q in ([0.1, 0.1, 0.1], [9.95, 5.0, 0.02], [5.0, 5.0, 5.0]):
        q = np.array(q)
        best = {}
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                for iz in (-1, 0, 1):
                    sh = q + np.array([ix, iy, iz]) * box
                    d = pos - sh
                    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
                    for pid, v in enumerate(d2):
                        if pid not in best or v < best[pid]:
                            best[pid] = v
        items = sorted(best.items(), key=lambda t: t[1])[:5]
        ids, r2 = c.query(q, 5)
        np.testing.assert_array_equal(ids[0], [t[0] for t in items])
        np.testing.assert_array_equal(r2[0], [t[1] for t in items])


def test_nonperiodic_knn_brute_force():
    rng = np.random.default_rng(7)
    pos = rng.random((500, 3))
    c = port.OracleCloud(pos, np.ones(500), [0, 1, 0, 1, 0, 1])
    q = rng.random((200, 3))
    ids, r2 = c.query(q, 4)
    for i in range(len(q)):
        d = q[i] - pos
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        order = np.argsort(d2, kind="stable")[:4]
        np.testing.assert_array_equal(np.sort(ids[i]), np.sort(order))
        np.testing.assert_array_equal(r2[i], d2[order])


def test_closed_forms_lattice():
    """test_degenerate.py:38-92 and test_traced_projection.py:130-138: a
    uniform field integrates to the ray length; sum(ds) telescopes."""
    pos, f, vol, bb = cubic_lattice()
    c = port.OracleCloud(pos, f, bb, vol=vol)
    s = np.array([r[0] for r in LATTICE_RAYS], float)
    e = np.array([r[1] for r in LATTICE_RAYS], float)
    vals, ids, se, sx, off = c.detailed(s, e)
    for i in range(len(s)):
        L = np.linalg.norm(e[i] - s[i])
        np.testing.assert_allclose(vals[i, 0], L, rtol=1e-6)
        np.testing.assert_allclose(np.sum(sx[off[i]:off[i + 1]] - se[off[i]:off[i + 1]]),
                                   L, rtol=1e-12)
        # consecutive segments share their boundary
        np.testing.assert_array_equal(se[off[i] + 1:off[i + 1]], sx[off[i]:off[i + 1] - 1])
        assert se[off[i]] == 0.0


def test_closed_forms_volume():
    """test_volume_render.py:20-58: transparent, uniform, saturated media."""
    rng = np.random.default_rng(42)
    pos = rng.random((500, 3)) * 100
    bb = [0., 100, 0, 100, 0, 100]
    s, e = np.array([[50, 50, 0.1]]), np.array([[50, 50, 99.9]])
    for rgba, check in (([1.0, 0.5, 0.25, 0.0], "zero"),
                        ([0.8, 0.4, 0.2, 0.01], "analytic"),
                        ([1.0, 0.5, 0.25, 1000.0], "saturated")):
        c = port.OracleCloud(pos, np.tile(rgba, (500, 1)), bb)
        v, _, se, sx, _ = c.detailed(s, e, "volume")
        L = np.sum(sx - se)
        if check == "zero":
            np.testing.assert_allclose(v[0], 0.0, atol=1e-15)
        elif check == "analytic":
            np.testing.assert_allclose(v[0], np.array(rgba[:3]) * (1 - np.exp(-0.01 * L)),
                                       rtol=1e-10)
        else:
            np.testing.assert_allclose(v[0], rgba[:3], rtol=1e-6)


def test_linearity_exact(galaxy):
    """test_multifield.py:16-62: the second field 2*dens integrates to exactly
    twice the first (scaling by 2 is exact in FP64)."""
    box = galaxy["box_size"]
    d = galaxy["dens"].astype(np.float64)
    c = port.OracleCloud(galaxy["pos"].astype(np.float64), np.column_stack([d, 2 * d]),
                         [0., box, 0., box, 0., box])
    s, e = _galaxy_rays(box, 32)
    v = c.project(s, e)
    np.testing.assert_array_equal(v[:, 1], 2 * v[:, 0])


def test_thread_count_invariance(galaxy):
    """test_openmp.py:99-129: results do not depend on the thread count."""
    box = galaxy["box_size"]
    args = (galaxy["pos"].astype(np.float64), galaxy["dens"], [0., box, 0., box, 0., box])
    s, e = _galaxy_rays(box, 48)
    v1 = port.OracleCloud(*args, nthreads=1).project(s, e)
    v4 = port.OracleCloud(*args, nthreads=4).project(s, e)
    np.testing.assert_array_equal(v1, v4)


def test_errors_and_warnings():
    with pytest.raises(ValueError, match="outside the bounding box"):
        port.OracleCloud(np.array([[5.0, 5, 5], [5, 5, 11.0]]), np.ones(2),
                         [0, 10, 0, 10, 0, 10], periodic=True)
    rng = np.random.default_rng(42)
    c = port.OracleCloud(rng.uniform(45, 55, (50, 3)), np.ones(50), [0, 100] * 3 and
                         [0, 100, 0, 100, 0, 100], periodic=True)
    assert any("60%" in w for w in c.warnings)
    c = port.OracleCloud(rng.random((50, 3)), np.ones(50), [0, 1, 0, 1, 0, 1])
    assert any("15% of longest box side" in w for w in c.warnings)
    with pytest.raises(ValueError, match="R, G, B"):
        c.project(np.zeros((1, 3)), np.ones((1, 3)), "volume")


# Each planted mistake (oracle/vt_oracle.c, vto_set_mutation) must make at
# least one pin fail; this shows the pins above are sharp enough.
MUTATIONS = {
    1: "equal r^2 inserted before earlier entries (tie order)",
    2: "child order flipped on exact ties",
    3: "distance summed in reverse dimension order",
    4: "periodic merge without de-duplication",
    5: "case 2 decided by the current cell only",
    6: "split-point dot product associated differently",
    7: "sum reduction as s_exit*f - s_enter*f",
    8: "periodic search over face images only",
    9: "zero-width step not skipped",
}


def _pins_hold():
    """The pins above in compact form; False as soon as one fails."""
    pos, f, vol, bb = cubic_lattice()
    a, b = _clouds(pos, f, bb, vol=vol)
    s = np.array([r[0] for r in LATTICE_RAYS], float)
    e = np.array([r[1] for r in LATTICE_RAYS], float)
    ra, rb = a.detailed(s, e), b.detailed(s, e)
    if not all(np.array_equal(x, y) for x, y in zip(ra, rb)):
        return False
    rng = np.random.default_rng(11)
    P = rng.random((3000, 3)) * 10
    F = np.exp(rng.standard_normal(3000))
    for periodic in (False, True):
        a, b = _clouds(P, F, [0, 10, 0, 10, 0, 10], periodic=periodic)
        gs, ge = grid.generate_projection_grid([0, 10], 12, [0, 10], None)
        ra = a.detailed(gs.reshape(-1, 3), ge.reshape(-1, 3))
        rb = b.detailed(gs.reshape(-1, 3), ge.reshape(-1, 3))
        if not all(np.array_equal(x, y) for x, y in zip(ra, rb)):
            return False
        q = rng.random((500, 3)) * 10
        q[:200] *= 0.05
        for k in (3, 5):
            ia, da = a.query(q, k)
            ib, db = b.query(q, k)
            if not (np.array_equal(ia, ib) and np.array_equal(da, db)):
                return False
    # a tiny periodic cloud: the same particle returns through several images
    rng2 = np.random.default_rng(5)
    P2 = rng2.random((12, 3))
    a, b = _clouds(P2, np.ones(12), [0, 1, 0, 1, 0, 1], periodic=True)
    q = rng2.random((50, 3))
    if not all(np.array_equal(x, y) for x, y in zip(a.query(q, 10), b.query(q, 10))):
        return False
    # lattice queries on exact ties (vertices, edge midpoints)
    pos, f, vol, bb = cubic_lattice()
    a, b = _clouds(pos, f, bb, vol=vol)
    q = np.array([[0.5, 0.5, 0.5], [1.5, 2.0, 2.5], [2.0, 2.5, 2.5], [3.5, 3.5, 0.5]])
    for k in (1, 3, 6):
        if not all(np.array_equal(x, y) for x, y in zip(a.query(q, k), b.query(q, k))):
            return False
    return True


ROTATED_VIEWS = [
    dict(extent=[10.0, 90.0], nres=24, bounds=[0.0, 100.0], center=[50.0, 50.0, 50.0],
         yaw=0.7, pitch=1.1, roll=0.3),
    dict(extent=[[0.0, 100.0], [20.0, 60.0]], nres=(9, 13), bounds=[100.0, 0.0],
         center=[50, 40, 60], yaw=2.5, pitch=0.2),
    dict(extent=[-3.0, 7.0], nres=5, bounds=[-1.0, 4.0], center=[1.0, 2.0, 3.0], proj="yz"),
    dict(extent=[0.0, 1.0], nres=(1, 3), bounds=[0.0, 1.0], center=[0.5, 0.5, 0.5],
         proj="zy"),
]


@pytest.mark.parametrize("view", range(len(ROTATED_VIEWS)))
def test_rotation_order_matches_numpy(view):
    """vto_rotate_about (the order a device generator uses) reproduces
    numpy.dot(p - c, R.T) + c of grid.generate_projection_grid bit for bit on
    this host; the unrotated grid is linspace/meshgrid, pinned separately."""
    v = dict(ROTATED_VIEWS[view])
    s_rot, e_rot = grid.generate_projection_grid(**v)
    base = dict(v)
    c = base.pop("center")
    s0, e0 = grid.generate_projection_grid(**base)  # same grid, no rotation
    if "proj" in v:
        yaw, pitch, roll, _ = grid._plane_angles(v["proj"])
    else:
        yaw, pitch, roll = v.get("yaw", 0.0), v.get("pitch", 0.0), v.get("roll", 0.0)
    R = grid.rotation_matrix(yaw, pitch, roll)
    np.testing.assert_array_equal(port.rotate_about(s0, R, c), s_rot.reshape(-1, 3))
    np.testing.assert_array_equal(port.rotate_about(e0, R, c), e_rot.reshape(-1, 3))
    # a plain sum in another order differs somewhere: the pin has teeth
    a = s0.reshape(-1, 3) - np.asarray(c, float)
    plain = (a[:, :1] * R[:, 0] + a[:, 1:2] * R[:, 1]) + a[:, 2:] * R[:, 2] + np.asarray(c, float)
    if view == 0:
        assert not np.array_equal(plain, s_rot.reshape(-1, 3))


@needs_ref
def test_pins_hold_unmutated():
    assert _pins_hold()


@needs_ref
@pytest.mark.parametrize("m", sorted(MUTATIONS))
def test_planted_mistake_is_caught(m):
    port.set_mutation(m)
    try:
        caught = not _pins_hold()
    except (RuntimeError, ValueError):
        caught = True
    finally:
        port.set_mutation(0)
    assert caught, f"mutation {m} ({MUTATIONS[m]}) passed every pin"
