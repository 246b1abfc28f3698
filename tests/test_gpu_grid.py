"""GPU tests of the ray grids built on the device (vt_grid_rays) and of the
batched grid projection that uses them.

* Unrotated grids (numpy.linspace pixel centres and meshgrid,
  vortrace/grid.py:35-83) are bit-identical to grid.generate_projection_grid.
* Rotated grids are bit-identical to the oracle's restatement of
  numpy.dot's evaluation order (oracle/vt_oracle.c vto_rotate_about), and to
  numpy itself wherever the host's BLAS evaluates in that order (checked
  here on the same inputs first).
* batch_grid_projection (device rays, one launch over every view) gives
  grid_projection's bits for each view.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import grid, synthetic  # noqa: E402

VIEWS = [
    dict(extent=[0.0, 100.0], nres=16, bounds=[0.0, 100.0]),
    dict(extent=[[5.0, 95.0], [10.0, 80.0]], nres=(7, 12), bounds=[100.0, 0.0]),
    dict(extent=[0, 100], nres=(1, 5), bounds=[0, 100]),
    dict(extent=[1e-3, 2e-3], nres=(3, 1), bounds=[0.0, 1.0]),
    dict(extent=[10.0, 90.0], nres=24, bounds=[0.0, 100.0], center=[50.0, 50.0, 50.0],
         yaw=0.7, pitch=1.1, roll=0.3),
    dict(extent=[[0.0, 100.0], [20.0, 60.0]], nres=(9, 13), bounds=[100.0, 0.0],
         center=[50, 40, 60], yaw=2.5, pitch=0.2),
    dict(extent=[0.0, 100.0], nres=8, bounds=[0.0, 100.0], center=[50.0, 50.0, 50.0],
         proj="yz"),
    dict(extent=[0.0, 100.0], nres=(6, 4), bounds=[0.0, 100.0], center=[50.0, 50.0, 50.0],
         proj="zy"),
    dict(extent=[0.0, 100.0], nres=(1, 1), bounds=[0.0, 100.0], center=[50.0, 50.0, 50.0],
         proj="xz"),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_grid(v):
    """The view's rays: numpy's unrotated grid, rotated by the oracle."""
    if v.get("center") is None:
        s, e = grid.generate_projection_grid(**v)
        return s.reshape(-1, 3), e.reshape(-1, 3)
    base = dict(v)
    c = base.pop("center")
    if "proj" in v:
        yaw, pitch, roll, _ = grid._plane_angles(v["proj"])
        base.pop("proj")
        # the plane's handedness flip applies before the rotation
        hand = grid._plane_angles(v["proj"])[3]
        if hand != np.sign(v["bounds"][1] - v["bounds"][0]):
            base["bounds"] = np.flip(v["bounds"])
    else:
        yaw, pitch, roll = v.get("yaw", 0.0), v.get("pitch", 0.0), v.get("roll", 0.0)
    for k in ("yaw", "pitch", "roll"):
        base.pop(k, None)
    s0, e0 = grid.generate_projection_grid(**base)
    R = grid.rotation_matrix(yaw, pitch, roll)
    return port.rotate_about(s0, R, c), port.rotate_about(e0, R, c)


def test_device_grids_match_numpy_and_oracle():
    _say(f"device ray grids: {len(VIEWS)} views")
    s, e, sizes = grid.generate_projection_grids_device(VIEWS)
    assert sizes == [int(np.prod(np.broadcast_to(v["nres"], 2))) for v in VIEWS]
    s, e = s.cpu().numpy(), e.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for k, v in enumerate(VIEWS):
        ds, de = s[off[k]:off[k + 1]], e[off[k]:off[k + 1]]
        os_, oe = _oracle_grid(v)
        np.testing.assert_array_equal(ds, os_, err_msg=f"view {k} start vs oracle")
        np.testing.assert_array_equal(de, oe, err_msg=f"view {k} end vs oracle")
        ns, ne = grid.generate_projection_grid(**v)
        ns, ne = ns.reshape(-1, 3), ne.reshape(-1, 3)
        if np.array_equal(os_, ns) and np.array_equal(oe, ne):
            np.testing.assert_array_equal(ds, ns, err_msg=f"view {k} start vs numpy")
            np.testing.assert_array_equal(de, ne, err_msg=f"view {k} end vs numpy")
        else:
            assert v.get("center") is not None  # only rotations depend on the host BLAS
            _say(f"  view {k}: host BLAS evaluates numpy.dot in another order; "
                 f"numpy comparison skipped")
    empty = grid.generate_projection_grids_device([])
    assert empty[0].shape == (0, 3) and empty[2] == []
    with pytest.raises(ValueError, match="nx, ny >= 1"):
        grid.generate_projection_grids_device([dict(extent=[0, 1], nres=(0, 3),
                                                    bounds=[0, 1])])


def test_batch_grid_projection_matches_grid_projection():
    _say("batched grid projection with device rays")
    pos, rho = synthetic.clustered_mesh(60_000, seed=8, n_clumps=20)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)
    # A periodic ray inside one cell raises the reference's bisection error
    # (src/ray.cpp:220-246), and so does any call holding one: the short view
    # does exactly what the oracle does, alone and with every view in one launch.
    short = [v for v in VIEWS if abs(float(v["bounds"][1]) - float(v["bounds"][0])) != 100.0]
    oc = port.OracleCloud(pos, rho, bb, periodic=True)
    for v in short:
        ds, de = _oracle_grid(v)
        try:
            want = oc.project(ds, de, "sum")[:, 0]
        except RuntimeError as ex:
            _say(f"  short view: the oracle raises {ex}")
            for views in ([v], VIEWS):
                with pytest.raises(RuntimeError, match="^" + str(ex) + "$"):
                    pc.batch_grid_projection(views, reduction="sum")
        else:
            np.testing.assert_array_equal(
                pc.batch_grid_projection([v], reduction="sum")[0].cpu().numpy().ravel(), want)
    views = [v for v in VIEWS if v not in short]
    out = pc.batch_grid_projection(views, reduction="sum")
    for k, (v, o) in enumerate(zip(views, out)):
        ds, de = _oracle_grid(v)
        ref = pc.projection(ds, de, reduction="sum")
        nres = v["nres"]
        nx, ny = (nres, nres) if isinstance(nres, int) else nres
        np.testing.assert_array_equal(o.cpu().numpy(), ref.reshape(nx, ny),
                                      err_msg=f"view {k}")
        if v.get("center") is None:
            np.testing.assert_array_equal(o.cpu().numpy(), pc.grid_projection(
                v["extent"], v["nres"], v["bounds"], reduction="sum"), err_msg=f"view {k}")
    assert pc.batch_grid_projection([], reduction="sum") == []
