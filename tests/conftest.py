"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU.

Fixture data live in tests/golden (see make_galaxy_fixture.py there); the
reference checkout is never read at test time.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def galaxy():
    """The reference's galaxy_interaction gas cells (tests/conftest.py:8-19 of
    the reference, converted from HDF5): pos float32 (7217, 3), dens float32,
    box_size float64."""
    d = np.load(os.path.join(GOLDEN, "galaxy_interaction_gas.npz"))
    return {"pos": d["pos"], "dens": d["dens"], "box_size": d["box_size"][()]}


@pytest.fixture(scope="session")
def galaxy_golden():
    return np.load(os.path.join(GOLDEN, "galaxy_interaction-proj.npy"))


def cubic_lattice(n=5):
    """make_cubic_lattice of the reference tests (e.g. test_degenerate.py:11-19)."""
    c = np.arange(n, dtype=np.float64)
    xx, yy, zz = np.meshgrid(c, c, c, indexing="ij")
    pos = np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])
    fields = np.ones(len(pos))
    vol = np.ones(len(pos))
    bb = [-0.5, n - 0.5, -0.5, n - 0.5, -0.5, n - 0.5]
    return pos, fields, vol, bb


# Rays of the reference's degenerate-geometry and traced tests
# (test_degenerate.py:38-92, test_vortrace.py:52-61, test_traced_projection.py).
LATTICE_RAYS = [
    ([-0.1, -0.1, -0.1], [4.1, 4.1, 4.1]),      # body diagonal through vertices
    ([0.5, 0.2, -0.1], [0.5, 0.2, 4.1]),        # along a face
    ([0.5, 0.5, -0.1], [0.5, 0.5, 4.1]),        # along an edge
    ([0.1, 0.1, 0.1], [0.9, 0.9, 0.9]),         # through one vertex
    ([0.50001, 0.50001, -0.1], [0.50001, 0.50001, 4.1]),  # near-degenerate
    ([2.0, 2.0, -0.1], [2.0, 2.0, 4.1]),        # through generators
    ([2.0, 2.0, -0.5], [2.0, 2.0, 4.5]),
    ([1.0, 2.0, -0.5], [1.0, 2.0, 4.5]),
    ([3.0, 3.0, -0.5], [3.0, 3.0, 4.5]),
    ([0.5, 0.5, 0.5], [0.6, 0.6, 0.6]),         # start on a vertex, same cell
    ([4.1, 4.1, 4.1], [-0.1, -0.1, -0.1]),      # reversed diagonal
    ([0.5, 2.0, 0.5], [3.5, 2.0, 3.5]),         # diagonal in a lattice plane
]


def _through(p, d, back=1.3, fwd=1.7):
    d = np.asarray(d, float) / np.linalg.norm(d)
    p = np.asarray(p, float)
    return (list(p - back * d), list(p + fwd * d))


# Oblique rays through Voronoi vertices and edge midpoints: case 1a with a
# different cell on each side (two points at one s, then a zero-width step).
LATTICE_RAYS += [
    _through([1.5, 1.5, 1.5], [1, 1, 0]),
    _through([1.5, 1.5, 1.5], [1, 2, 0]),
    _through([1.5, 1.5, 1.5], [1, 2, 3]),
    _through([2.5, 1.5, 2.0], [1, 1, 1]),
    _through([2.0, 2.5, 1.5], [0, 1, 2]),
]


def grid_cloud(box, spacing, field_val=1.0):
    """_make_grid_cloud of the reference (test_periodic.py:10-18) as arrays."""
    c = np.arange(spacing / 2.0, box, spacing)
    xx, yy, zz = np.meshgrid(c, c, c, indexing="ij")
    pos = np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])
    return pos, np.full(len(pos), field_val), [0.0, box, 0.0, box, 0.0, box]
