"""CPU tests of bench.py's contract: the algorithmic-bytes count behind the
roofline fraction, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n_points, n_f, per_crossing", [
    (100_000, 1, 1384),    # 100k uniform, sum
    (2_000_000, 1, 1640),  # 2M clustered, sum
    (2_000_000, 4, 1664),  # 2M clustered, volume (RGBA fields)
    (4_000_000, 1, 1704),  # 4M clustered, sum
    (8_000_000, 1, 1768),  # 8M clustered, sum
])
def test_algorithmic_bytes_per_crossing(n_points, n_f, per_crossing):
    """Per crossing: one field gather (8 n_f) and two root-to-leaf queries of
    32 B nodes over D = ceil(log2(N / 10)) levels plus a 10-point leaf of
    24 B generators; the figures are the ones the measurement plan states for
    each mesh size."""
    C = 1_000_000
    assert bench.algorithmic_bytes(0, C, n_points, n_f, 1) == per_crossing * C


def test_algorithmic_bytes_per_ray_and_clustered_launch():
    # per ray: start and end (48 B) and the n_out outputs
    assert bench.algorithmic_bytes(1000, 0, 2_000_000, 1, 3) == 1000 * (48 + 24)
    # the 2M clustered 512^2 forward launch: 262,144 rays, 33,471,991 crossings
    b = bench.algorithmic_bytes(262_144, 33_471_991, 2_000_000, 1, 1)
    assert b == 262_144 * 56 + 33_471_991 * 1640
    assert abs(b / 1e9 - 54.909) < 1e-3


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
