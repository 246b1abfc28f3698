"""Host-side checks that need no GPU: the C-ABI library, the product's tree
builder, ray grids and the Python API's validation."""
import ctypes as C
import os
import re
import warnings

import numpy as np
import pytest

from conftest import ROOT, cubic_lattice
from oracle import port, ref
from vortrace_amd import _lib, grid, synthetic
from vortrace_amd.core import PointCloud
from vortrace_amd.vortrace import ProjectionCloud


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "vortrace_hip.h")) as f:
        text = f.read()
    declared = set(re.findall(r"\b(vt_[a-z_0-9]+)\s*\(", text))
    declared.discard("vt_warning_fn")
    assert declared, "no declarations found"
    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)


def _product_tree(pos, fields, bb, vol=None, periodic=False):
    pc = PointCloud()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pc.loadPoints(pos, fields, bb, vol, periodic)
    pc.build_host_tree()
    return pc


def _reference_tree(pos, fields, bb, vol=None, periodic=False):
    if ref.available():
        c = ref.RefCloud(pos, fields, bb, vol=vol, periodic=periodic)
        return ref.parse_tree_bytes(c.tree_bytes()), c.orig_ids
    c = port.OracleCloud(pos, fields, bb, vol=vol, periodic=periodic)
    return c.tree(), c.orig_ids


@pytest.mark.parametrize("case", ["galaxy", "lattice", "uniform", "clustered",
                                  "duplicates"])
def test_tree_matches_nanoflann(case, galaxy):
    """The device tree's slot order, cut dims and divlow/divhigh equal the
    reference's nanoflann index (include/nanoflann.hpp:858-1003)."""
    vol, periodic = None, False
    if case == "galaxy":
        box = galaxy["box_size"]
        pos, f, bb = galaxy["pos"].astype(np.float64), galaxy["dens"], [0., box] * 3
        bb = [0., box, 0., box, 0., box]
    elif case == "lattice":
        pos, f, vol, bb = cubic_lattice(7)
    elif case == "uniform":
        rng = np.random.default_rng(3)
        pos, f, bb = rng.random((50000, 3)) * 100, np.ones(50000), [0, 100, 0, 100, 0, 100]
    elif case == "clustered":
        pos, f = synthetic.clustered_mesh(80000, seed=9, workers=4)
        bb, periodic = [0, 100, 0, 100, 0, 100], True
    else:  # many exactly repeated coordinates
        rng = np.random.default_rng(4)
        pos = np.round(rng.random((20000, 3)) * 8) / 8
        f, bb = np.ones(20000), [0, 1, 0, 1, 0, 1]
    pc = _product_tree(pos, f, bb, vol, periodic)
    (vind, bbox, nodes), orig = _reference_tree(pos, f, bb, vol, periodic)
    v2, b2, n2 = pc.tree_export()
    np.testing.assert_array_equal(pc.get_orig_ids(), orig)
    np.testing.assert_array_equal(v2, vind)
    np.testing.assert_array_equal(b2, bbox)
    assert n2 == nodes


@pytest.mark.parametrize("case", ["uniform", "lattice"])
def test_tree_bytes(case):
    """saveTreeToBytes writes nanoflann's index dump of the same tree; a dump
    (ours, or the reference's own where oracle/_ref is built) loads back into
    the same tree; malformed dumps raise ValueError."""
    if case == "uniform":
        rng = np.random.default_rng(8)
        pos, f, bb, vol = rng.random((30000, 3)) * 100, np.ones(30000), [0, 100] * 3, None
        bb = [0, 100, 0, 100, 0, 100]
    else:
        pos, f, vol, bb = cubic_lattice(6)
    pc = _product_tree(pos, f, bb, vol)
    data = pc.saveTreeToBytes()
    vind, bbox, nodes = ref.parse_tree_bytes(data)
    v2, b2, n2 = pc.tree_export()
    np.testing.assert_array_equal(vind, v2)
    np.testing.assert_array_equal(bbox, b2)
    assert nodes == n2
    dumps = [data]
    if ref.available():
        rc = ref.RefCloud(pos, f, bb, vol=vol)
        rbytes = rc.tree_bytes()
        rv, rb, rn = ref.parse_tree_bytes(rbytes)
        np.testing.assert_array_equal(rv, vind)
        np.testing.assert_array_equal(rb, bbox)
        assert rn == nodes
        assert len(rbytes) == len(data)
        dumps.append(rbytes)
    for d in dumps:
        q = _product_tree(pos, f, bb, vol)
        q.load_host_tree_bytes(d)
        v3, b3, n3 = q.tree_export()
        np.testing.assert_array_equal(v3, v2)
        np.testing.assert_array_equal(b3, b2)
        assert n3 == n2
        assert q.saveTreeToBytes() == data
    q = _product_tree(pos, f, bb, vol)
    for bad, msg in [(data[:100], "truncated"), (data[:-40], "truncated"),
                     (b"\x07" + data[1:], "points"),
                     (data[:8] + b"\x02" + data[9:], "dimension")]:
        with pytest.raises(ValueError, match=msg):
            q.load_host_tree_bytes(bad)
    empty = PointCloud()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        empty.loadPoints(np.zeros((0, 3)), np.zeros(0), bb)
    with pytest.raises(RuntimeError, match="no points loaded"):
        empty.load_host_tree_bytes(data)
    with pytest.raises(RuntimeError, match="has not been built"):
        empty.saveTreeToBytes()


def test_cloud_getters_and_filtering():
    pos = np.array([[0.5, 0.5, 0.5], [0.9, 0.9, 0.9], [10.0, 10.0, 10.0]])
    pc = PointCloud()
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        pc.loadPoints(pos, np.array([1.0, 2.0, 3.0]), [0, 1, 0, 1, 0, 1])
    assert any("15% of longest box side" in str(x.message) for x in w)
    assert pc.get_npart() == 2 and pc.get_nfields() == 1
    assert not pc.get_periodic() and not pc.get_tree_built()
    np.testing.assert_array_equal(pc.get_orig_ids(), [0, 1])
    assert pc.get_pad() == pytest.approx(0.15)
    assert pc.get_subbox() == [0, 1, 0, 1, 0, 1]
    assert pc.get_field(1, 0) == 2.0 and pc.get_pt(1) == [0.9, 0.9, 0.9]


def test_periodic_validation():
    with pytest.raises(ValueError, match="outside the bounding box"):
        ProjectionCloud(np.array([[5.0, 5, 5], [-0.1, 5, 5]]), np.ones(2),
                        boundbox=[0, 10, 0, 10, 0, 10], periodic=True,
                        _skip_build_tree=True)
    rng = np.random.default_rng(42)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        ProjectionCloud(rng.uniform(45.0, 55.0, size=(50, 3)), np.ones(50),
                        boundbox=[0, 100, 0, 100, 0, 100], periodic=True,
                        _skip_build_tree=True)
    assert any("60%" in str(x.message) for x in w)


def test_vol_padding():
    """test_padding.py:10-45: pad = 10 * r_max(vol) keeps nearby points."""
    rng = np.random.default_rng(42)
    pos = np.vstack([rng.uniform(20, 80, size=(100, 3)), [[5.0, 50.0, 50.0]]])
    vol = np.full(len(pos), (4.0 / 3.0) * np.pi * 5.0 ** 3)
    pc = ProjectionCloud(pos, np.ones(len(pos)), boundbox=[10.0, 90.0] * 3 and
                         [10.0, 90.0, 10.0, 90.0, 10.0, 90.0], vol=vol,
                         _skip_build_tree=True)
    assert len(pc.orig_ids) == len(pos)


def test_reduction_validation_messages():
    rng = np.random.default_rng(1)
    pos = rng.random((100, 3)) * 10
    bb = [0, 10, 0, 10, 0, 10]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pc = ProjectionCloud(pos, rng.random(100), boundbox=bb, _skip_build_tree=True)
        with pytest.raises(ValueError, match="Unknown reduction"):
            pc.projection(np.zeros((1, 3)), np.ones((1, 3)), reduction="bogus")
        with pytest.raises(ValueError, match="4 fields"):
            pc.grid_projection([2, 8], 8, [0, 10], None, reduction="volume")
        f = np.column_stack([rng.random(100) * 2, rng.random(100), rng.random(100),
                             rng.random(100) * 0.1])
        pc = ProjectionCloud(pos, f, boundbox=bb, _skip_build_tree=True)
        with pytest.raises(ValueError, match=r"R, G, B.*\[0, 1\]"):
            pc.grid_projection([2, 8], 8, [0, 10], None, reduction="volume")
        f[:, 0] = 0.5
        f[:, 3] = -0.1
        pc = ProjectionCloud(pos, f, boundbox=bb, _skip_build_tree=True)
        with pytest.raises(ValueError, match="alpha >= 0"):
            pc.traced_projection([5, 5, 0.1], [5, 5, 9.9], reduction="volume")
        pc = ProjectionCloud(pos, rng.random(100), boundbox=bb, _skip_build_tree=True)
        with pytest.raises(RuntimeError, match="no valid tree"):
            pc.projection(np.zeros((1, 3)), np.ones((1, 3)))
        with pytest.raises(ValueError):
            pc.traced_projection(np.zeros((3, 3)), np.zeros((4, 3)))
        with pytest.raises(ValueError):
            pc.traced_projection(np.zeros(4), np.zeros(4))
        with pytest.raises(AttributeError, match="traced_projection"):
            pc.single_projection(np.zeros(3), np.ones(3))


def test_grid_values_exact():
    """test_grid.py:12-42 (exact equality)."""
    g0 = grid.generate_base_grid([0, 1], 64)
    g1 = grid.generate_base_grid([[0, 1], [0, 1]], (64, 64))
    np.testing.assert_equal(g0, g1)
    assert g0[0][0][0] == 0.5 / 64. and g0[-1][-1][0] == 63.5 / 64.
    g = grid.generate_base_grid([[0, 1], [0, 3]], (64, 128))
    assert g[0][0][1] == 0.5 / (128. / 3.) and g[-1][-1][1] == 127.5 / (128 / 3.)
    g = grid.generate_base_grid([3, 4], 64)
    assert g[0][0][0] == 3 + 0.5 / 64. and g[-1][-1][1] == 3 + 63.5 / 64.


@pytest.mark.parametrize("proj,expected", [
    ("xy", [[[0.25, 0.25, 0.], [0.25, 0.75, 0.]], [[0.75, 0.25, 0.], [0.75, 0.75, 0.]]]),
    ("yz", [[[0., 0.25, 0.25], [0., 0.25, 0.75]], [[0., 0.75, 0.25], [0., 0.75, 0.75]]]),
    ("zx", [[[0.25, 0., 0.25], [0.75, 0., 0.25]], [[0.25, 0., 0.75], [0.75, 0., 0.75]]]),
    ("yx", [[[0.25, 0.25, 0.], [0.75, 0.25, 0.]], [[0.25, 0.75, 0.], [0.75, 0.75, 0.]]]),
    ("zy", [[[0., 0.25, 0.25], [0., 0.75, 0.25]], [[0., 0.25, 0.75], [0., 0.75, 0.75]]]),
    ("xz", [[[0.25, 0., 0.25], [0.25, 0., 0.75]], [[0.75, 0., 0.25], [0.75, 0., 0.75]]]),
])
def test_projection_planes(proj, expected):
    """test_grid.py:45-117."""
    gs, _ = grid.generate_projection_grid([0, 1], 2, [0, 1], [0.5, 0.5, 0.5], proj=proj)
    np.testing.assert_almost_equal(gs, np.array(expected))


def test_synthetic_is_seeded_and_bounded():
    p1, r1 = synthetic.clustered_mesh(30000, seed=5, workers=2)
    p2, r2 = synthetic.clustered_mesh(30000, seed=5, workers=2)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(r1, r2)
    assert p1.min() >= 0.0 and p1.max() < 100.0
    from scipy.spatial import cKDTree
    d, _ = cKDTree(p1, boxsize=100.0).query(p1, k=2)
    assert d[:, 1].min() >= 1e-3


def _tile_order_loops(n, row):
    """The 8x8-tile fetch order written as loops: bands of 8 rows, tiles of 8
    columns (fewer at the right edge), rows of a tile, columns of a row; the
    rays past the last whole band in index order."""
    out = []
    nbands = n // (8 * row) if row else 0
    for b in range(nbands):
        for t0 in range(0, row, 8):
            for r in range(8):
                for c in range(t0, min(t0 + 8, row)):
                    out.append((8 * b + r) * row + c)
    return out + list(range(len(out), n))


@pytest.mark.parametrize("n,row", [(0, 5), (7, 0), (64, 8), (481, 13), (960, 24),
                                   (279, 31), (120, 2000), (256, 1)])
def test_tiled_slots(n, row):
    from vortrace_amd.autograd import tiled_slots
    assert tiled_slots(n, row).tolist() == _tile_order_loops(n, row)


@pytest.mark.parametrize("n,row,parts,block", [(1, 0, 8, 4), (37, 0, 8, 128),
                                               (1000, 0, 8, 64), (1000, 0, 1, 256),
                                               (4096, 64, 8, 1 << 15), (4800, 24, 8, 192),
                                               (999, 13, 8, 320)])
def test_longest_first_order(n, row, parts, block):
    """longest_first: a permutation of [0, n); cut the fetch order (tiles of
    `row`-ray rows) in runs of 64 slots: every block of `block` slots of a ray
    partition holds its own runs, whole, by descending largest count, slot
    order among equals."""
    import torch
    from vortrace_amd.autograd import longest_first
    from vortrace_amd import tuning
    counts = torch.from_numpy(np.random.default_rng(n).integers(0, 40, n))
    old = tuning.get_tuning()
    try:
        tuning.set_tuning(ray_parts=parts)
        order = longest_first(counts, row, block)
    finally:
        tuning.set_tuning(**old)
    o = order.numpy().astype(np.int64)
    assert order.dtype == torch.int32
    assert np.array_equal(np.sort(o), np.arange(n))
    fetch = _tile_order_loops(n, row)
    c = counts.numpy()
    psize = -(-n // parts)
    for p in range(parts):
        for lo in range(p * psize, min(n, (p + 1) * psize), block):
            hi = min(n, (p + 1) * psize, lo + block)
            runs = {}
            for sl in range(lo, hi):
                runs.setdefault((sl - lo) // 64, []).append(sl)
            ranked = sorted(runs, key=lambda k: (-max(c[fetch[sl]] for sl in runs[k]), k))
            expect = [fetch[sl] for k in ranked for sl in runs[k]]
            assert list(o[lo:hi]) == expect



def test_default_tuning_and_environment_overrides():
    """Launch parameters at load: the defaults with no VT_* variable set, and
    each variable overriding its own parameter (out-of-range values ignored)."""
    import json
    import subprocess
    import sys
    code = ("import json, vortrace_amd as vt; "
            "print(json.dumps(vt.tuning.get_tuning()))")

    def load(**env_vars):
        env = {k: v for k, v in os.environ.items() if not k.startswith("VT_")}
        env.update(env_vars)
        out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, check=True,
                             capture_output=True, text=True, timeout=120).stdout
        return json.loads(out.strip().splitlines()[-1])

    assert load() == {"walk_batch": 32, "leaf_batch": 32, "ray_parts": 8, "help_events": 128}
    assert load(VT_HELP_EVENTS="0")["help_events"] == 0
    assert load(VT_HELP_EVENTS="7", VT_WALK_BATCH="16") == {
        "walk_batch": 16, "leaf_batch": 32, "ray_parts": 8, "help_events": 7}
    assert load(VT_HELP_EVENTS="129", VT_RAY_PARTS="1")["help_events"] == 128
    assert load(VT_RAY_PARTS="1")["ray_parts"] == 1
