"""Launch parameters of the ray kernels and an autotuner for them.

The ray engine (vortrace_amd/csrc/engine.h) runs node steps every iteration
and groups leaf steps and walk steps: a wave runs its leaf steps once
``leaf_batch`` of its lanes are at a leaf, and its walk steps once
``walk_batch`` lanes wait for one.  Rays are taken from ``ray_parts``
partitions (8: one per XCD, for L2 locality; or 1).  Once a lane has no ray
left to take, it walks an interval of a busy lane's ray of its wave, recording
up to ``help_events`` boundaries of it (0: off).  Results never depend on
these values, only speed.  The kd-tree's leaf size stays the reference's 10
(src/pointcloud.cpp:204): the order of equidistant neighbours, and with it the
crossing sequences, depends on it.
"""
from __future__ import annotations

import ctypes as C
import itertools

import torch

from . import _lib


def get_tuning() -> dict:
    t = _lib.Tuning()
    _lib.check(_lib.lib().vt_get_tuning(C.byref(t)))
    return {"walk_batch": t.walk_batch, "leaf_batch": t.leaf_batch, "ray_parts": t.ray_parts,
            "help_events": t.help_events}


def set_tuning(walk_batch: int | None = None, leaf_batch: int | None = None,
               ray_parts: int | None = None, help_events: int | None = None) -> dict:
    """Sets the given parameters (process-wide); returns the new settings."""
    cur = get_tuning()
    t = _lib.Tuning(int(walk_batch if walk_batch is not None else cur["walk_batch"]),
                    int(leaf_batch if leaf_batch is not None else cur["leaf_batch"]),
                    int(ray_parts if ray_parts is not None else cur["ray_parts"]),
                    int(help_events if help_events is not None else cur["help_events"]))
    _lib.check(_lib.lib().vt_set_tuning(C.byref(t)))
    return get_tuning()


_CACHE: dict = {}


def autotune(pc, pos_start, pos_end, *, reduction: str = "integrate", stride: int = 4,
             walk_batches=(16, 24, 32, 48), leaf_batches=(16, 24, 32, 40),
             ray_parts=(1, 8), help_events=(128,), repeats: int = 2,
             cache: bool = True) -> dict:
    """Times every combination on every `stride`-th ray of (pos_start,
    pos_end) through ``pc`` (a ProjectionCloud) and keeps the fastest.
    Returns {"best": settings, "kernel_ms": best time, "results": [...],
    "cached": bool}.  With `cache`, a second call for the same cloud, ray
    count, reduction and search space applies the stored best settings
    without timing again."""
    key = (id(pc._cloud), len(pos_start), str(reduction), int(stride), tuple(walk_batches),
           tuple(leaf_batches), tuple(ray_parts), tuple(help_events))
    if cache and key in _CACHE:
        hit = _CACHE[key]
        set_tuning(**hit["best"])
        return dict(hit, cached=True)
    dev = pc._cloud._device
    s = torch.as_tensor(pos_start, dtype=torch.float64).reshape(-1, 3)[::stride]
    e = torch.as_tensor(pos_end, dtype=torch.float64).reshape(-1, 3)[::stride]
    s, e = s.to(dev).contiguous(), e.to(dev).contiguous()
    pc.projection_tensor(s, e, reduction=reduction)  # warm-up
    results = []
    for wb, lb, rp, he in itertools.product(walk_batches, leaf_batches, ray_parts, help_events):
        set_tuning(wb, lb, rp, he)
        ms = []
        for _ in range(repeats):
            pc.projection_tensor(s, e, reduction=reduction)
            ms.append(_lib.last_call_stats().kernel_ms)
        results.append({"walk_batch": wb, "leaf_batch": lb, "ray_parts": rp,
                        "help_events": he, "kernel_ms": min(ms)})
    best = min(results, key=lambda r: r["kernel_ms"])
    set_tuning(best["walk_batch"], best["leaf_batch"], best["ray_parts"], best["help_events"])
    out = {"best": get_tuning(), "kernel_ms": best["kernel_ms"], "results": results}
    if cache:
        _CACHE[key] = out
    return dict(out, cached=False)
