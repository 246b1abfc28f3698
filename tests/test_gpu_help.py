"""GPU tests of intervals walked by other lanes (vt_tuning.help_events).

A lane with no ray left to take walks the bottom interval of a busy lane's
pending list (its wave's), and the ray's own lane replays the interval's
boundaries.  The walk of an interval depends only on its two end points and
the ray (src/ray.cpp:250-424), so every setting must give the bits of the
plain walk:

* help_events 0 (off), 1 and 3 (almost every interval is handed back
  unwalked and walked again by its own lane), 128 (the default);
* sum with one and with three fields, max, min, and the traced projection
  (whose count pass helps and whose fill pass does not: a count that differs
  fails the fill);
* uniform, clustered periodic and degenerate lattice clouds, with fewer rays
  than resident lanes (most lanes idle from the start, so intervals are given
  in every wave) and with more.
The statistics show that intervals were given and handed back.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import _lib, synthetic  # noqa: E402
from conftest import LATTICE_RAYS, cubic_lattice  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture()
def tuning():
    saved = vt.tuning.get_tuning()
    yield
    vt.tuning.set_tuning(**saved)


def _rays(n, seed, box=100.0):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.0, box, (n, 3))
    e = rng.uniform(0.0, box, (n, 3))
    return s, e


def _grid(npix, box=100.0):
    s, e = vt.grid.generate_projection_grid([0.0, box], npix, [0.0, box], None)
    return s.reshape(-1, 3), e.reshape(-1, 3)


def _project_all(pc, s, e, modes):
    out = {}
    for m in modes:
        out[m] = pc.projection(s, e, reduction=m)
        st = _lib.last_call_stats()
        out[m + ":intervals"] = (st.intervals, st.intervals_back)
    vals, ids, sv, ds, off = pc.traced_projection(s, e, flatten=True)
    out["traced"] = (vals, ids, sv, ds, off)
    return out


def _check_settings(pc, s, e, modes, name):
    res = {}
    for he in (0, 1, 3, 128):
        vt.tuning.set_tuning(help_events=he)
        _say(f"{name}: help_events {he}, {len(s)} rays")
        res[he] = _project_all(pc, s, e, modes)
    for he in (1, 3, 128):
        for m in modes:
            np.testing.assert_array_equal(res[he][m], res[0][m], err_msg=f"{name} {m} {he}")
        for a, b in zip(res[he]["traced"], res[0]["traced"]):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} traced {he}")
    for m in modes:
        assert res[0][m + ":intervals"] == (0, 0)
    return res


def test_uniform_few_rays(tuning):
    pos, rho = synthetic.uniform_mesh(30_000, seed=11)
    pc = vt.ProjectionCloud(pos, rho, boundbox=[0, 100, 0, 100, 0, 100])
    s, e = _rays(3000, 12)
    res = _check_settings(pc, s, e, ["integrate", "max", "min"], "uniform")
    given, back = res[128]["integrate:intervals"]
    assert given > 0, "no interval was given to another lane"
    given1, back1 = res[1]["integrate:intervals"]
    assert given1 > 0 and back1 > 0, "no interval came back at help_events 1"
    oc = port.OracleCloud(pos, rho, [0, 100, 0, 100, 0, 100], periodic=False)
    np.testing.assert_array_equal(res[128]["integrate"], oc.project(s, e, "sum")[:, 0])


def test_clustered_periodic_three_fields(tuning):
    pos, rho = synthetic.clustered_mesh(60_000, seed=13, n_clumps=30)
    f3 = np.stack([rho, rho * rho, np.sqrt(rho)], axis=1)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, f3, boundbox=bb, periodic=True)
    s, e = _grid(48)
    res = _check_settings(pc, s, e, ["integrate", "max"], "clustered nf=3")
    assert res[128]["integrate:intervals"][0] > 0
    oc = port.OracleCloud(pos, f3, bb, periodic=True)
    np.testing.assert_array_equal(res[128]["integrate"], oc.project(s, e, "sum"))


def test_clustered_periodic_many_rays(tuning):
    """More rays than resident lanes: intervals are given only at the end."""
    pos, rho = synthetic.clustered_mesh(200_000, seed=14, n_clumps=60)
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)
    lanes = pc._cloud.lanes(pc._validate_reduction("integrate"))
    nx = 512
    ny = -(-lanes * 3 // 2 // nx)
    x, y = np.meshgrid((np.arange(nx) + 0.5) * (100.0 / nx), (np.arange(ny) + 0.5) * (100.0 / ny))
    s = np.column_stack([x.ravel(), y.ravel(), np.zeros(nx * ny)])
    e = np.column_stack([x.ravel(), y.ravel(), np.full(nx * ny, 100.0)])
    res = {}
    for he in (0, 128):
        vt.tuning.set_tuning(help_events=he)
        _say(f"many rays: help_events {he}, {len(s)} rays")
        res[he] = pc.projection(s, e, reduction="integrate")
        if he:
            assert _lib.last_call_stats().intervals > 0
    np.testing.assert_array_equal(res[128], res[0])


def test_lattice_degenerate_rays(tuning):
    """Zero-width steps and perpendicular cycling inside given intervals."""
    pos, f, vol, bb = cubic_lattice()
    pc = vt.ProjectionCloud(pos, f, boundbox=bb, vol=vol)
    s = np.array([r[0] for r in LATTICE_RAYS], dtype=np.float64)
    e = np.array([r[1] for r in LATTICE_RAYS], dtype=np.float64)
    _check_settings(pc, s, e, ["integrate", "max", "min"], "lattice")


@pytest.mark.parametrize("seed", [21, 22, 23, 24, 25, 26])
def test_random_meshes_and_rays(tuning, seed):
    """Random clustered meshes (periodic and not) and random chords: helping
    on (any pair given) and off give the same bits, and both equal the oracle."""
    rng = np.random.default_rng(seed)
    periodic = bool(seed % 2)
    n = int(rng.integers(5_000, 40_000))
    pos, rho = synthetic.clustered_mesh(n, seed=seed, n_clumps=int(rng.integers(3, 20)))
    bb = [0, 100, 0, 100, 0, 100]
    pc = vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=periodic)
    nr = int(rng.integers(500, 3000))
    if periodic:  # axis-aligned rays across the box (the reference's periodic use)
        xy = rng.uniform(0.0, 100.0, (nr, 2))
        ax = rng.integers(0, 3)
        s = np.insert(xy, ax, 0.0, axis=1)
        e = np.insert(xy, ax, 100.0, axis=1)
    else:
        s = rng.uniform(5.0, 95.0, (nr, 3))
        e = rng.uniform(5.0, 95.0, (nr, 3))
    out = {}
    for he in (0, 128):
        vt.tuning.set_tuning(help_events=he)
        _say(f"random mesh seed {seed}: {n} points, {nr} rays, periodic {periodic}, "
             f"help_events {he}")
        out[he] = (pc.projection(s, e, reduction="integrate"),
                   pc.traced_projection(s, e, flatten=True))
    np.testing.assert_array_equal(out[128][0], out[0][0])
    for a, b in zip(out[128][1], out[0][1]):
        np.testing.assert_array_equal(a, b)
    oc = port.OracleCloud(pos, rho, bb, periodic=periodic)
    np.testing.assert_array_equal(out[128][0], oc.project(s, e, "sum")[:, 0])
