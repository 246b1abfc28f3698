"""GPU tests of the batched multi-view launch and of the launch-parameter
autotuner.

* Views of different sizes and directions projected in one launch give the
  same bits as each view on its own and as the oracle (oracle/vt_oracle.c,
  pinned to the reference) -- the multi-view configuration at a size the
  oracle finishes in seconds.
* The launch parameters change speed only: every setting, and the one the
  autotuner keeps, gives the same bits.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import port

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import synthetic  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def clustered():
    pos, rho = synthetic.clustered_mesh(120_000, seed=4, n_clumps=40)
    bb = [0, 100, 0, 100, 0, 100]
    return pos, rho, bb, vt.ProjectionCloud(pos, rho, boundbox=bb, periodic=True)


def test_batched_views_match_single_and_oracle(clustered):
    pos, rho, bb, pc = clustered
    views = synthetic.camera_rays(44, 6, sizes=(8, 16, 24, 32))
    _say(f"batched views: {[len(s) for s, _ in views]} rays")
    batched = pc.batch_projection(views, reduction="sum")
    oc = port.OracleCloud(pos, rho, bb, periodic=True)
    for (s, e), b in zip(views, batched):
        np.testing.assert_array_equal(b, pc.projection(s, e, reduction="sum"))
        np.testing.assert_array_equal(b, oc.project(s, e, "sum")[:, 0])
    assert pc.batch_projection([], reduction="sum") == []


def test_tuning_changes_speed_only(clustered):
    pos, rho, bb, pc = clustered
    views = synthetic.camera_rays(45, 3, sizes=(32, 48))
    s = np.concatenate([v[0] for v in views])
    e = np.concatenate([v[1] for v in views])
    saved = vt.tuning.get_tuning()
    try:
        ref = pc.projection(s, e, reduction="sum")
        for wb, lb, rp, he in [(1, 1, 1, 0), (64, 64, 8, 2), (8, 40, 1, 128)]:
            vt.tuning.set_tuning(wb, lb, rp, he)
            np.testing.assert_array_equal(pc.projection(s, e, reduction="sum"), ref,
                                          err_msg=f"tuning {wb} {lb} {rp} {he}")
        space = dict(reduction="sum", stride=2, walk_batches=(16, 32), leaf_batches=(16, 32),
                     ray_parts=(1, 8), help_events=(0, 128), repeats=1)
        out = vt.tuning.autotune(pc, s, e, **space)
        assert set(out["best"]) == {"walk_batch", "leaf_batch", "ray_parts", "help_events"}
        assert out["best"]["help_events"] in (0, 128) and not out["cached"]
        assert len(out["results"]) == 16
        assert out["kernel_ms"] == min(r["kernel_ms"] for r in out["results"])
        np.testing.assert_array_equal(pc.projection(s, e, reduction="sum"), ref)
        # the same cloud, rays and space again: the stored settings, no timing
        vt.tuning.set_tuning(8, 40, 1, 3)
        again = vt.tuning.autotune(pc, s, e, **space)
        assert again["cached"] and again["best"] == out["best"]
        assert vt.tuning.get_tuning() == out["best"]
        np.testing.assert_array_equal(pc.projection(s, e, reduction="sum"), ref)
        with pytest.raises(ValueError, match="tuning out of range"):
            vt.tuning.set_tuning(walk_batch=0)
        with pytest.raises(ValueError, match="help_events in"):
            vt.tuning.set_tuning(help_events=129)
    finally:
        vt.tuning.set_tuning(**saved)
