"""GPU tests of the walk's failure paths and of K_EXPAND on ties.

* The reference's runtime errors raised inside the walk come out of the GPU
  kernels with the reference's exception type and text: a periodic ray inside
  one cell (src/ray.cpp:220-246) and the iteration cap (src/ray.cpp:260-261,
  a ray across 100,000 slab cells).  The kernels record the error and exit
  normally; the call reports it.
* K_EXPAND with every one of the 100 nearest generators equidistant
  (src/ray.cpp:320-338): integer points on a sphere (exact ties) around the
  split point.  Without periodic images the 100-NN result and its tie order
  are nanoflann's, so the crossings equal the oracle's and the reference's;
  when periodic images are merged the tie order would follow libstdc++'s
  introsort, and the GPU refuses with a RuntimeError instead.

Texts are compared with the reference's own build (oracle/_ref) where it is
present, else with the oracle (pinned to it).
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port, ref

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cpu_cloud(*args, **kw):
    return ref.RefCloud(*args, **kw) if ref.available() else port.OracleCloud(*args, **kw)


def _cpu_error(cloud, s, e, mode="sum"):
    with pytest.raises(RuntimeError) as ex:
        cloud.project(s, e, mode)
    return str(ex.value)


def test_periodic_ray_inside_one_cell():
    rng = np.random.default_rng(1)
    pos = rng.random((1000, 3)) * 10
    f = np.ones(1000)
    bb = [0, 10, 0, 10, 0, 10]
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, periodic=True)
    s = (pos[0] + [1e-4, 0, 0])[None]
    e = (pos[0] + [2e-4, 0, 0])[None]
    want = _cpu_error(_cpu_cloud(pos, f, bb, periodic=True), s, e)
    _say(f"  periodic single-cell ray: the reference raises {want!r}")
    assert want == "Could not find intermediate cell by bisection along periodic ray"
    with pytest.raises(RuntimeError) as ex:
        pc.projection(s, e)
    assert str(ex.value) == want
    # one such ray fails the whole call, as the reference's makeProjection
    rays_s = np.concatenate([s, [[0.5, 0.5, 0.0]] * 64])
    rays_e = np.concatenate([e, [[0.5, 0.5, 10.0]] * 64])
    with pytest.raises(RuntimeError, match="^" + want + "$"):
        pc.projection(rays_s, rays_e)
    with pytest.raises(RuntimeError, match="^" + want + "$"):
        pc.traced_projection(s[0], e[0])
    # the cloud stays usable
    np.testing.assert_array_equal(pc.projection(rays_s[1:], rays_e[1:]),
                                  port.OracleCloud(pos, f, bb, periodic=True).project(
                                      rays_s[1:], rays_e[1:])[:, 0])


def test_iteration_cap():
    n = 100_000
    pos = np.zeros((n, 3))
    pos[:, 2] = np.arange(n, dtype=np.float64)
    f = np.ones(n)
    vol = np.ones(n)
    bb = [-1, 1, -1, 1, -0.5, n - 0.5]
    s = np.array([[0.1, 0.2, 0.25]])
    e = np.array([[0.1, 0.2, n - 1.25]])
    want = _cpu_error(_cpu_cloud(pos, f, bb, vol=vol), s, e)
    _say(f"  {n} slab cells along one ray: the reference raises {want!r}")
    assert want == "Ray walk exceeded maximum iteration count"
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    for fn in (pc.projection, pc.traced_projection):
        with pytest.raises(RuntimeError) as ex:
            fn(s, e)
        assert str(ex.value) == want
    # a shorter ray through the same cells stays under the cap
    s2, e2 = np.array([[0.1, 0.2, 0.25]]), np.array([[0.1, 0.2, 20000.25]])
    np.testing.assert_array_equal(pc.projection(s2, e2),
                                  port.OracleCloud(pos, f, bb, vol=vol).project(s2, e2)[:, 0])


def _sphere_points(r):
    """Integer points with x^2 + y^2 + z^2 = r^2 (exact ties at the centre)."""
    pts = set()
    n = r * r
    for x in range(-r, r + 1):
        for y in range(-r, r + 1):
            z2 = n - x * x - y * y
            if z2 < 0:
                continue
            z = int(round(z2 ** 0.5))
            if z * z == z2:
                pts.add((x, y, z))
                pts.add((x, y, -z))
    return np.array(sorted(pts), dtype=np.float64)


def _sphere_cloud(seed, r=45, keep=118):
    """The two poles and `keep` other points of the sphere: a ray along the
    axis splits the poles' cells at the centre, where all 120 points are at
    r^2 exactly, so the 3-NN holds no pole and the 100-NN is all ties."""
    pts = _sphere_points(r)
    poles = [i for i, p in enumerate(pts) if p[0] == 0 and p[1] == 0]
    others = [i for i in range(len(pts)) if i not in poles]
    rng = np.random.default_rng(seed)
    sel = np.sort(np.concatenate([poles, rng.choice(others, keep, replace=False)]))
    p = pts[sel]
    return p, 1.0 + 0.01 * np.arange(len(p))


@pytest.mark.parametrize("seed", [0, 3, 7])
def test_expand_all_equidistant_non_periodic(seed):
    r = 45
    pos, f = _sphere_cloud(seed, r)
    bb = [-r - 1, r + 1] * 3
    vol = np.ones(len(pos))
    s = np.array([[0.0, 0.0, -r - 0.5], [0.0, 0.0, r + 0.5]])
    e = np.array([[0.0, 0.0, r + 0.5], [0.0, 0.0, -r - 0.5]])
    oc = port.OracleCloud(pos, f, bb, vol=vol)
    ov, oids, ose, osx, ooff = oc.detailed(s, e)
    if ref.available():
        rv, rids, rse, rsx, roff = ref.RefCloud(pos, f, bb, vol=vol).detailed(s, e)
        assert np.array_equal(rids, oids) and np.array_equal(rsx, osx)
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    v, ids, se, sx, off = vt.Projection(s, e).segments_tensor(pc._cloud, vt.ReductionMode.Sum)
    st = vt._lib.last_call_stats()
    _say(f"  sphere seed {seed}: {ooff[-1]} segments, {st.slow_queries} K_EXPAND searches")
    assert st.slow_queries > 0
    np.testing.assert_array_equal(off.cpu().numpy(), ooff)
    np.testing.assert_array_equal(ids.cpu().numpy(), oids)
    np.testing.assert_array_equal(se.cpu().numpy(), ose)
    np.testing.assert_array_equal(sx.cpu().numpy(), osx)
    np.testing.assert_array_equal(v.cpu().numpy(), ov)
    np.testing.assert_array_equal(pc.projection(s, e), ov[:, 0])


def test_expand_all_equidistant_periodic():
    """Far from the box faces no image is searched: the result is the
    reference's.  In a box whose faces are within the 100-NN radius the
    images merge and the GPU refuses (documented divergence)."""
    r = 45
    pos, f = _sphere_cloud(0, r)
    s = np.array([[0.0, 0.0, -r - 0.5]])
    e = np.array([[0.0, 0.0, r + 0.5]])
    big = 400.0
    bb = [0.0, big] * 3
    shift = np.array([big / 2] * 3)
    oc = port.OracleCloud(pos + shift, f, bb, periodic=True)
    pc = vt.ProjectionCloud(pos + shift, f, boundbox=bb, periodic=True)
    np.testing.assert_array_equal(pc.projection(s + shift, e + shift),
                                  oc.project(s + shift, e + shift)[:, 0])
    assert vt._lib.last_call_stats().slow_queries > 0
    # sphere centred at (r, r, r) in a box of side 100: from the centre the
    # +x, +y and +z images lie at r^2 exactly (searched and merged)
    bb2 = [0.0, 100.0] * 3
    p2 = pos + r
    s2 = np.array([[r, r, 0.25]])
    e2 = np.array([[r, r, 2 * r - 0.25]])
    pc2 = vt.ProjectionCloud(p2, f, boundbox=bb2, periodic=True)
    with pytest.raises(RuntimeError) as ex:
        pc2.projection(s2, e2)
    try:
        port.OracleCloud(p2, f, bb2, periodic=True).project(s2, e2)
    except RuntimeError as oex:
        assert str(ex.value) == str(oex)
    else:
        assert "equidistant" in str(ex.value) and "periodic images" in str(ex.value)
    _say(f"  sphere in a tight periodic box: {ex.value}")
