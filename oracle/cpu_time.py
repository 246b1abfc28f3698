"""Times a CPU projection for bench.py's cpu_baseline leg (test and
measurement infrastructure, never product).

    python -m oracle.cpu_time IN.npz OUT.npz KIND MODE PERIODIC REPEATS

IN.npz holds pos, fields, start, end; KIND is "reference" (oracle/_ref, the
reference's own C++ core) or "port" (the oracle's C restatement).  Runs in
its own process so that the OpenMP settings of its environment
(OMP_NUM_THREADS, OMP_PROC_BIND, OMP_PLACES) take effect when the OpenMP
runtime starts.  The cloud is built first (tree resident, as on the GPU
side); each timed run is one makeProjection over every ray.  OUT.npz gets
the values of the last run, the seconds of each run and the crossing count
(from one untimed detailed projection).
"""
from __future__ import annotations

import sys
import time

import numpy as np


def main(argv):
    src, dst, kind, mode, periodic, repeats = argv
    d = np.load(src)
    pos, fields, start, end = d["pos"], d["fields"], d["start"], d["end"]
    if kind == "reference":
        from oracle import ref as impl
        cloud = impl.RefCloud(pos, fields, [0, 100, 0, 100, 0, 100], periodic=periodic == "1")
    else:
        from oracle import port as impl
        cloud = impl.OracleCloud(pos, fields, [0, 100, 0, 100, 0, 100],
                                 periodic=periodic == "1")
    cloud.project(start[:64], end[:64], mode)  # warm-up: threads started, pages touched
    secs = []
    vals = None
    for _ in range(int(repeats)):
        t0 = time.perf_counter()
        vals = cloud.project(start, end, mode)
        secs.append(time.perf_counter() - t0)
    off = cloud.detailed(start, end, mode)[4]
    np.savez(dst, values=vals, seconds=np.array(secs), crossings=np.int64(off[-1]))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
