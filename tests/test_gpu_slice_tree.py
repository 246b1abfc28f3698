"""GPU tests of slice (src/slice.cpp:7-60) and of projecting through a tree
loaded from nanoflann index bytes (src/pointcloud.cpp:243-281).

* slice: every pixel's field values equal the oracle's (oracle/vt_oracle.c)
  1-NN of the reference's pixel coordinates, bit for bit; out-of-box extents
  raise the reference's RuntimeError.
* A tree loaded from the reference's own index dump (oracle/_ref, where it
  is built) or from this build's dump projects bit-identically to the oracle.
"""
import sys

import numpy as np
import pytest

from oracle import port, ref

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import grid, synthetic  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("periodic", [False, True])
def test_slice_matches_oracle(periodic):
    _say(f"slice periodic={periodic}")
    pos, rho = synthetic.uniform_mesh(20000, seed=21)
    fields = np.column_stack([rho, 2.0 * rho + 1.0])
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, fields, boundbox=bb, periodic=periodic)
    extent, (nx, ny), depth = [5.0, 95.0, 10.0, 90.0], (17, 23), 37.5
    got = pc.slice(extent, (nx, ny), depth)
    assert got.shape == (nx, ny, 2)
    # src/slice.cpp:26-47: x = xmin + dx * i, dx = (xmax - xmin) / (nx - 1)
    dx = (extent[1] - extent[0]) / (nx - 1)
    dy = (extent[3] - extent[2]) / (ny - 1)
    pts = np.array([[extent[0] + dx * i, extent[2] + dy * j, depth]
                    for i in range(nx) for j in range(ny)])
    oc = port.OracleCloud(pos, fields, bb, periodic=periodic)
    ids, _ = oc.query(pts, 1)
    expect = fields[oc.orig_ids][ids[:, 0]].reshape(nx, ny, 2)
    np.testing.assert_array_equal(got, expect)
    one = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    assert one.slice(extent, 9, depth).shape == (9, 9)
    with pytest.raises(RuntimeError, match="Slice extent out of bounds"):
        pc.slice([-1.0, 50.0, 0.0, 50.0], 8, 10.0)
    with pytest.raises(RuntimeError, match="Slice extent out of bounds"):
        pc.slice([0.0, 50.0, 0.0, 50.0], 8, 101.0)


def test_projection_through_loaded_tree():
    _say("projection through loaded tree bytes")
    pos, rho = synthetic.uniform_mesh(30000, seed=22)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)
    s, e = grid.generate_projection_grid([0.0, 100.0], 24, [0.0, 100.0], None)
    s, e = s.reshape(-1, 3), e.reshape(-1, 3)
    expect = port.OracleCloud(pos, rho, bb, periodic=True).project(s, e, "sum")[:, 0]
    dumps = [pc.save_tree_bytes()]
    if ref.available():
        dumps.append(ref.RefCloud(pos, rho, bb, periodic=True).tree_bytes())
    for data in dumps:
        q = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)
        q.load_tree_bytes(data)
        np.testing.assert_array_equal(q.projection(s, e, reduction="sum"), expect)
        assert q.save_tree_bytes() == dumps[0]


def test_tree_file_and_ray_class(tmp_path):
    """PointCloud.saveTree / loadTree with files (src/pointcloud.cpp:214-240)
    and the Ray class (src/bindings.cpp:277-307) against the reference's
    build (the oracle where it is absent)."""
    _say("tree files and Ray")
    pos, rho = synthetic.uniform_mesh(20000, seed=23)
    f4 = np.column_stack([0.2 + 0.5 * np.minimum(rho / 10, 1), np.full(len(rho), 0.3),
                          np.full(len(rho), 0.6), 0.02 * rho])
    bb = [0, 100, 0, 100, 0, 100]
    cpu = (ref.RefCloud if ref.available() else port.OracleCloud)(pos, f4, bb, periodic=True)
    pc = vt.PointCloud()
    pc.loadPoints(pos, f4, bb, periodic=True)
    with pytest.raises(RuntimeError, match="Cannot save tree: tree has not been built"):
        pc.saveTree(str(tmp_path / "t.bin"))
    pc.buildTree()
    path = tmp_path / "tree.bin"
    pc.saveTree(str(path))
    if ref.available():
        # the same index (the reference leaves its node unions' padding bytes
        # uninitialised, so compare the parsed dumps)
        mine, theirs = ref.parse_tree_bytes(path.read_bytes()), ref.parse_tree_bytes(
            cpu.tree_bytes())
        assert np.array_equal(mine[0], theirs[0]) and np.array_equal(mine[1], theirs[1])
        assert mine[2] == theirs[2]
    with pytest.raises(RuntimeError, match="Cannot open file for writing"):
        pc.saveTree(str(tmp_path / "missing" / "t.bin"))
    q = vt.PointCloud()
    q.loadPoints(pos, f4, bb, periodic=True)
    with pytest.raises(RuntimeError, match="Cannot open file for reading"):
        q.loadTree(str(tmp_path / "nothing.bin"))
    q.loadTree(str(path))
    assert q.saveTreeToBytes() == path.read_bytes()
    empty = vt.PointCloud()
    empty.loadPoints(np.zeros((0, 3)), np.zeros(0), bb, periodic=True)
    with pytest.raises(RuntimeError, match="Cannot load tree: no points loaded"):
        empty.loadTree(str(path))
    start, end = np.array([3.0, 97.0, 0.5]), np.array([88.0, 12.0, 99.5])
    for mode, name, get in ((vt.ReductionMode.Sum, "sum", "get_col"),
                            (vt.ReductionMode.Max, "max", "get_max_val"),
                            (vt.ReductionMode.Min, "min", "get_min_val"),
                            (vt.ReductionMode.VolumeRender, "volume", "get_vol_render_val")):
        r = vt.Ray(start, end)
        assert getattr(r, get)() == [] and r.get_dens_col() == 0.0
        r.integrate(q, mode)
        vals, ids, se, sx, off = cpu.detailed(start[None], end[None], name)
        if name == "volume":
            np.testing.assert_allclose(getattr(r, get)(), vals[0], rtol=1e-12, atol=0)
        else:
            assert getattr(r, get)() == [float(x) for x in vals[0]]
        segs = r.get_segments()
        assert [s[0] for s in segs] == ids.tolist()
        assert [s[1] for s in segs] == se.tolist() and [s[2] for s in segs] == sx.tolist()
        assert [s[3] for s in segs] == (sx - se).tolist()
        if name == "sum":
            assert r.get_dens_col() == vals[0][0]
    assert vt.Ray(start, end).walk(q) == segs
