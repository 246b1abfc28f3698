"""Benchmark: cell-crossings/s and rays/s of the forward projection on one GPU.

Default workload (one step = one full image):
  2,000,000-point clustered periodic mesh (box 100, seed 2), 512^2 rays along
  +z over [0, 100], reduction 'sum', FP64 throughout.
Inputs (tree, fields, rays) are resident on the device before the timed
region; tree build and host->device copies are reported separately.

Other workloads (--workload): uniform100k (100k uniform points, 128^2),
clustered2m_volume (volume rendering at 1024^2), multiview4m (4M points, 64
cameras of random direction, extent and size in ONE batched launch), fit8m
(8M points, 16 views at 512^2: forward + adjoint of a squared loss, with a
central-difference check on 64 cells) and fit1m_volume (the volume adjoint on
a 1M mesh at 256^2, with a central-difference check).

Prints ONE JSON line.  Extra keys: rays_per_s, kernel_ms, fwd_adjoint (the
same rays, forward plus the adjoint of a squared loss through torch.autograd),
roofline (HBM bound,
algorithmic bytes per launch / average kernel time from HIP events recorded on
the kernel's stream), cpu_baseline (the reference's OpenMP C++ core, built from
its sources into oracle/_ref, timed on this host on a sample of the same rays;
checked bit-identical to the GPU values on that sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E vendor peak
METRIC = ("cell-crossings/s + rays/s, fwd vs fwd+adjoint, 2M clustered cells, "
          "512² image")

WORKLOADS = {
    "uniform100k": dict(desc="100k uniform points in [0,100)^3 (seed 1), non-periodic, "
                    "128^2 z-rays over [0,100], sum",
               n=100_000, npix=128, periodic=False, mode="sum"),
    "clustered2m": dict(desc="2M clustered periodic mesh (box 100, seed 2), 512^2 z-rays "
                    "over [0,100], sum",
               n=2_000_000, npix=512, periodic=True, mode="sum"),
    "clustered2m_volume": dict(desc="2M clustered periodic mesh (seed 2), RGBA transfer function, "
                    "1024^2 z-rays over [0,100], volume rendering",
               n=2_000_000, npix=1024, periodic=True, mode="volume"),
    "multiview4m": dict(desc="4M clustered periodic mesh (seed 4), 64 cameras (seed 40): "
                    "isotropic directions through the box centre, extents U[10,100], "
                    "npix from {64,128,256,512,1024}; one batched launch, sum",
               n=4_000_000, seed=4, periodic=True, mode="sum", kind="views", views=64,
               cam_seed=40, sizes=(64, 128, 256, 512, 1024), extent=(10.0, 100.0)),
    "fit8m": dict(desc="8M clustered periodic mesh (seed 5), 16 views at 512^2 (cameras seed "
                  "50, extent 100); L = sum (proj(rho) - target)^2, target = "
                  "proj(rho exp(0.1 N(0,1))) (seed 55); forward + adjoint by torch.autograd",
               n=8_000_000, seed=5, periodic=True, mode="sum", kind="fit", views=16,
               cam_seed=50, sizes=(512,), extent=(100.0, 100.0)),
    "fit1m_volume": dict(desc="1M clustered periodic mesh (seed 6), RGBA transfer function, "
                         "256^2 z-rays over [0,100]; L = sum (proj(rgba) - target)^2, target "
                         "with alpha scaled by exp(0.1 N(0,1)) (seed 55); volume forward + "
                         "adjoint by torch.autograd",
               n=1_000_000, seed=6, npix=256, periodic=True, mode="volume", kind="fit"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    """Mesh, fields and views [(start, end), ...] of a workload."""
    from vortrace_amd import grid, synthetic
    w = WORKLOADS[name]
    t0 = time.perf_counter()
    if name == "uniform100k":
        pos, fields = synthetic.uniform_mesh(w["n"], seed=1)
    else:
        pos, fields = synthetic.clustered_mesh(w["n"], seed=w.get("seed", 2))
    log(f"[bench] mesh {name}: {len(pos)} points in {time.perf_counter() - t0:.1f}s "
        f"(sha {synthetic.digest(pos, fields)})")
    params = None
    if "views" in w:
        params = synthetic.camera_views(w["cam_seed"], w["views"], sizes=w["sizes"],
                                        extent_range=w["extent"])
        views = []
        for v in params:
            s, e = grid.generate_projection_grid(**v)
            views.append((s.reshape(-1, 3), e.reshape(-1, 3)))
    else:
        s, e = grid.generate_projection_grid([0.0, 100.0], w["npix"], [0.0, 100.0], None)
        views = [(s.reshape(-1, 3), e.reshape(-1, 3))]
    return pos, fields, views, params


def volume_fields(pos, rho):
    """RGBA fields of the volume workloads (periodic box of side 100): kappa
    such that the median opacity of a 64^2 z-probe is 0.5, then the
    transfer function.  Returns (rgba (N, 4), kappa)."""
    import torch
    import vortrace_amd as vt
    from vortrace_amd import synthetic
    pc_probe = vt.ProjectionCloud(pos, rho, boundbox=[0, 100, 0, 100, 0, 100], periodic=True)
    ps, pe = vt.grid.generate_projection_grid([0.0, 100.0], 64, [0.0, 100.0], None)
    _, ids, se, sx, off = pc_probe.segments_tensor(
        torch.from_numpy(ps.reshape(-1, 3)), torch.from_numpy(pe.reshape(-1, 3)))
    kappa = synthetic.kappa_for_median_opacity(
        rho, ids.cpu().numpy(), (sx - se).cpu().numpy(), off.cpu().numpy())
    del pc_probe
    return synthetic.rgba_transfer(rho, kappa), kappa


def algorithmic_bytes(R, C, n_points, n_f, n_out):
    """R(48 + 8 n_out) + C(8 n_f) + 2C(32 D + 240), D = ceil(log2(N/10)):
    ray endpoints and outputs, one field gather per crossing, and two kNN
    queries per crossing, each one root-to-leaf descent of 32-byte nodes plus
    one 10-point leaf of 24-byte generators."""
    D = math.ceil(math.log2(n_points / 10.0))
    return R * (48 + 8 * n_out) + C * (8 * n_f) + 2 * C * (32 * D + 240)


def adjoint_bytes(R, C, n_points, n_f, n_out, sort_passes):
    """Algorithmic bytes the one-walk adjoint adds to the forward's: per
    segment the 32-byte record written by the forward and read back (64), its
    (slot, ds) written in ray order (12), the contribution step (ds read 8,
    contributions written 8 n_c; for volume also slot 4 and the cell's 4
    fields 32 read, twice), the stable radix sort of (slot, index) pairs
    (16 bytes per pass), run bounds (4) and the per-cell sums (index 4 and
    contribution 8 n_c); per ray the count written, scanned and read (24)
    and the image gradient read (8 n_out); per cell the gradient written
    (8 n_f) with its run bounds and slot id (12).  n_c = n_f for sum, 4 for
    volume."""
    vol = n_out == 3
    n_c = 4 if vol else n_f
    per_seg = (64 + 12 + 8 + 8 * n_c + (2 * (4 + 32) if vol else 0) + 16 * sort_passes + 4
               + 4 + 8 * n_c)
    return C * per_seg + R * (24 + 8 * n_out) + n_points * (8 * n_f + 12)


def sort_passes(n_points):
    """Digit passes of the adjoint's radix sort of slots (8-bit digits)."""
    return max(1, math.ceil(max(1, math.ceil(math.log2(max(2, n_points)))) / 8))


def source_digest():
    """SHA-256 (16 hex digits) of the HIP library's sources: ties a recorded
    profile to the kernels it measured."""
    import hashlib
    d = os.path.join(ROOT, "vortrace_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".h", ".cpp")):
            with open(os.path.join(d, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE runs, scaled by a copy-kernel calibration) and the L2 hit rate
    (TCC_HIT / (TCC_HIT + TCC_MISS)) of the same kernel, as recorded in
    profiles/traffic_<workload>.json, with the kernel time of the profiled run
    and whether the recorded sources are the current ones."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return {"hbm_bytes_per_launch": d.get("hbm_bytes_per_launch"),
            "l2_hit_rate": d.get("l2_hit_rate"),
            "profiled_kernel_ms": d.get("kernel_ms"),
            "profiled_source_digest": d.get("source_digest"),
            "matches_current_source": d.get("source_digest") == source_digest()}


def stream_copy_gbs(gib=2.0, repeats=5):
    """Measured device copy bandwidth (read + write bytes / time) of a
    `gib`-GiB FP64 copy, best of `repeats`, timed with HIP events."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    best = None
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    torch.cuda.empty_cache()
    return 2 * 8 * n / (best * 1e-3) / 1e9


def host_cpu_info():
    """CPU model, sockets, cores per socket and threads per core (lscpu), and
    the CPUs this process may run on."""
    info = {"model": None, "sockets": None, "cores_per_socket": None,
            "threads_per_core": None}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, ValueError):
        pass
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    if info["sockets"] and info["cores_per_socket"]:
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    info["smt"] = (info["threads_per_core"] or 1) > 1
    return info


def cpu_baseline(pos, fields, start, end, periodic, mode, frac, gpu_vals):
    """The reference's C++ core (oracle/_ref; the oracle's restatement where
    the reference build is absent) on a strided sample of the rays, in a
    child process with OMP_NUM_THREADS = the physical cores available (at most
    the box's CPU share, OMP_NUM_THREADS of this environment),
    OMP_PROC_BIND=close and OMP_PLACES=cores; median of 3 timed runs."""
    import subprocess
    import tempfile
    from oracle import ref
    kind = "reference" if ref.available() else "port"
    stride = max(1, int(round(1.0 / frac)))
    idx = np.arange(0, len(start), stride)
    cpu = host_cpu_info()
    share = int(os.environ.get("OMP_NUM_THREADS", cpu["affinity_cpus"]))
    cores = cpu.get("physical_cores") or cpu["affinity_cpus"]
    threads = max(1, min(share, cores, cpu["affinity_cpus"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as tmp:
        np.savez(os.path.join(tmp, "in.npz"), pos=pos, fields=fields, start=start[idx],
                 end=end[idx])
        cmd = [sys.executable, "-m", "oracle.cpu_time", os.path.join(tmp, "in.npz"),
               os.path.join(tmp, "out.npz"), kind, mode, "1" if periodic else "0", "3"]
        r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True)
        if r.returncode != 0:
            log(f"[bench] cpu baseline failed: {r.stderr[-2000:]}")
            return None
        res = np.load(os.path.join(tmp, "out.npz"))
        vals, times, crossings = res["values"], res["seconds"], int(res["crossings"])
    dt = float(np.median(times))
    gv = gpu_vals[idx]
    same = bool(np.array_equal(vals.reshape(gv.shape), gv))
    close = bool(np.allclose(vals.reshape(gv.shape), gv, rtol=1e-12, atol=0))
    log(f"[bench] cpu {kind}: {len(idx)} rays, {crossings} crossings, median "
        f"{dt:.2f}s of {len(times)} ({threads} threads on {cpu['model']}); values "
        f"bit-identical: {same}")
    return {"value": crossings / dt, "unit": "cell-crossings/s", "cores": threads,
            "kind": kind,
            "sample": f"every {stride}th ray of the same rays ({len(idx)} rays, "
                      f"{crossings} crossings), "
                      + ("the reference's C++ core built from its sources "
                         "(makeProjection, OpenMP dynamic,256)" if kind == "reference" else
                         "the oracle's C restatement (OpenMP dynamic,256)")
                      + f"; median of {len(times)} runs",
            "seconds": [float(t) for t in times],
            "rays_per_s": len(idx) / dt,
            "cpu": cpu,
            "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close",
                    "OMP_PLACES": "cores"},
            "values_bit_identical_to_gpu": same,
            "values_within_1e-12": close}


def fit_target(pc, fields, d_start, d_end, mode):
    """Target image of the fit: proj(rho * exp(0.1 N(0,1) per cell)) for sum;
    for volume the alpha column scaled by exp(0.1 N(0,1)) (RGBA stays valid).
    Noise seed 55."""
    from vortrace_amd import synthetic
    z = synthetic.normal(np.random.default_rng(55), len(fields))
    if mode == "sum":
        noisy = fields * np.exp(0.1 * z)
    else:
        noisy = fields.copy()
        noisy[:, 3] *= np.exp(0.1 * z)
    pc.set_fields(noisy)
    target = pc.projection_tensor(d_start, d_end, reduction=mode).clone()
    pc.set_fields(fields)
    return target[:, 0].contiguous() if mode == "sum" else target


def fwd_adjoint(pc, fields, d_start, d_end, args, C, R, mode="sum"):
    """Forward plus adjoint on the same rays: one step evaluates
    L = sum over pixels of (proj(fields) - target)^2 and dL/dfields through
    torch.autograd (projection kernel forward, adjoint kernels backward).
    Returns (report, gradient, target)."""
    import torch
    import vortrace_amd as vt
    dev = d_start.device
    target = fit_target(pc, fields, d_start, d_end, mode)
    f_t = torch.from_numpy(np.ascontiguousarray(fields, dtype=np.float64)).to(dev)

    split = {"forward_ms": [], "backward_ms": []}

    def step(rs=d_start, re=d_end):
        f = f_t.clone().requires_grad_(True)
        t0 = time.perf_counter()
        p = vt.autograd.project(pc, f, rs, re, mode)
        loss = ((p - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        split["forward_ms"].append(1e3 * (t1 - t0))
        split["backward_ms"].append(1e3 * (time.perf_counter() - t1))
        return f.grad, pc._forward_stats, pc._adjoint_stats

    first = None  # the first pass over these rays (record pool from a sample)
    if args.warmup > 0:
        # The process's first autograd step, on copies of the rays (a ray set
        # of its own): one-time costs of torch's autograd, the loss kernels and
        # the first record pool, kept out of the first pass over the step's rays.
        step(d_start.clone(), d_end.clone())
        first = {"process_first_step_ms": split["forward_ms"][0] + split["backward_ms"][0]}
        split["forward_ms"].clear()
        split["backward_ms"].clear()
    for i in range(args.warmup):
        step()
        if i == 0:
            first.update({"ms": split["forward_ms"][0] + split["backward_ms"][0],
                          "adjoint_path": pc._adjoint_path})
        log(f"[bench] fwd+adjoint warmup {i + 1}/{args.warmup} done")
    split["forward_ms"].clear()
    split["backward_ms"].clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fms, kms = [], []
    g = None
    for i in range(args.steps):
        g, fst, ast = step()
        fms.append(fst.kernel_ms)
        kms.append(ast.kernel_ms)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    nz = int(torch.count_nonzero(g.reshape(len(fields), -1).abs().amax(dim=1)).item())
    fwd_ms, bwd_ms = float(np.mean(split["forward_ms"])), float(np.mean(split["backward_ms"]))
    log(f"[bench] fwd+adjoint {args.steps} steps: {ms:.2f} ms/step (forward {fwd_ms:.2f}, "
        f"backward {bwd_ms:.2f}), forward kernels {float(np.mean(fms)):.2f} ms, adjoint "
        f"kernels {float(np.mean(kms)):.2f} ms, {nz} cells with a gradient")
    rep = {"ms_per_step": ms, "crossings_per_s": C / (ms * 1e-3), "rays_per_s": R / (ms * 1e-3),
           "forward_kernel_ms": float(np.mean(fms)),
           "adjoint_kernel_ms": float(np.mean(kms)), "forward_ms": fwd_ms,
           "backward_ms": bwd_ms,
           "step_ms_median": float(np.median(np.add(split["forward_ms"], split["backward_ms"]))),
           "cells_with_gradient": nz, "first_step": first,
           "step": "L = sum (proj(fields) - target)^2 and dL/dfields by torch.autograd "
                   "(forward kernel recording its segments + adjoint kernels over the "
                   "records), FP64"}
    return rep, g, target


def fd_check(pc, fields, grad, target, d_start, d_end, mode, ncheck=64, seed=56):
    """Central differences of the fit loss on `ncheck` random cells with a
    nonzero gradient (one random channel each for RGBA).  L changes only
    through the rays that cross the cell, and is evaluated as
    L(+h) - L(-h) = sum over those rays of (p+ - p-)(p+ + p- - 2 t), which
    avoids the cancellation of the whole image's sum.  Steps: 1e-3 relative
    for sum and for R, G, B (L is quadratic in each of them, so the difference
    is exact up to rounding) and 1e-4 relative for alpha (truncation
    <= (h ds)^2 / 6 relative).  E, the worst-case rounding of the
    difference (sum over the cell's rays of 2 n eps (|p+| + |p-|)
    |p+ + p- - 2t| / 2h, n = the ray's segments), must be below
    1e-7 max(|fd|, |g|) for a cell to be checked; cells whose influence on
    the image is below that resolution are drawn again.  Checked cells must
    satisfy |fd - g| <= 1e-6 max(|fd|, |g|).  The cells drawn again are
    checked too, against what their resolution allows:
    |fd - g| <= E + 1e-6 max(|fd|, |g|)."""
    import torch
    import vortrace_amd as vt
    dev = d_start.device
    rng = np.random.default_rng(seed)
    g = grad.detach().reshape(len(fields), -1)
    nf = g.shape[1]
    f0 = torch.from_numpy(np.ascontiguousarray(fields, dtype=np.float64)).to(dev)
    f0 = f0.reshape(len(fields), -1)
    cand = torch.nonzero(g.abs().amax(dim=1) > 0).flatten().cpu().numpy()
    order = np.argsort(rng.random(len(cand)), kind="stable")
    picks = []
    for c in cand[order]:
        ch = min(int(nf * rng.random()), nf - 1)
        v = float(f0[c, ch])
        h = (1e-4 if mode == "volume" and ch == 3 else 1e-3) * abs(v)
        if h == 0.0 or (mode == "volume" and ch < 3 and (v - h < 0.0 or v + h > 1.0)):
            continue
        picks.append((int(c), ch, v, h))
        if len(picks) == 8 * ncheck:
            break
    mark = torch.zeros(len(fields), dtype=torch.bool, device=dev)
    mark[torch.tensor([p[0] for p in picks], dtype=torch.int64, device=dev)] = True
    _, ids, _, _, off = pc.segments_tensor(d_start, d_end, reduction=mode)
    hit = torch.nonzero(mark[ids]).flatten()
    hit_cell = ids[hit]
    hit_ray = torch.searchsorted(off, hit, right=True) - 1
    nseg = off[1:] - off[:-1]
    del ids, mark
    tgt = target.reshape(target.shape[0], -1)
    eps = float(np.finfo(np.float64).eps)
    worst, strict, ok, nrays, checked, unresolved = 0.0, 0, True, [], 0, 0
    low_ok, low_worst = 0, 0.0  # cells below the resolution: |fd - g| / (E + 1e-6 scale)
    for c, ch, v, h in picks:
        if checked == ncheck:
            break
        rays = torch.unique(hit_ray[hit_cell == c])
        nrays.append(int(rays.numel()))
        s_sub, e_sub, t_sub = d_start[rays].contiguous(), d_end[rays].contiguous(), tgt[rays]
        n_sub = nseg[rays].to(torch.float64).reshape(-1, 1)

        def proj(val):
            f = f0.clone()
            f[c, ch] = val
            pc.set_fields(f.reshape(fields.shape))
            return pc.projection_tensor(s_sub, e_sub, reduction=mode)

        pp, pm = proj(v + h), proj(v - h)
        step = (v + h) - (v - h)
        w = (pp + pm - 2.0 * t_sub)
        fd = float(((pp - pm) * w).sum()) / step
        rnd = float((2.0 * n_sub * eps * (pp.abs() + pm.abs()) * w.abs()).sum()) / step
        gv = float(g[c, ch])
        scale = max(abs(fd), abs(gv))
        if not rnd <= 1e-7 * scale:
            unresolved += 1
            r = abs(fd - gv) / (rnd + 1e-6 * scale) if rnd + scale > 0 else 0.0
            low_ok += r <= 1.0
            low_worst = max(low_worst, r)
            continue
        checked += 1
        rel = abs(fd - gv) / scale
        strict += rel <= 1e-6
        worst = max(worst, rel)
    pc.set_fields(fields)
    ok = checked == ncheck and strict == checked and low_ok == unresolved
    log(f"[bench] central differences on {checked} cells ({min(nrays)}-{max(nrays)} rays "
        f"each; {unresolved} drawn cells below the rounding resolution redrawn): "
        f"{strict} within 1e-6 relative (max {worst:.2e}); redrawn cells within their "
        f"rounding bound: {low_ok}/{unresolved} (max err/bound {low_worst:.2e})")
    return {"cells": checked, "within_1e-6": int(strict), "max_rel_err": worst,
            "all_within_1e-6": bool(ok), "redrawn_below_resolution": unresolved,
            "redrawn_within_rounding_bound": int(low_ok),
            "redrawn_max_err_over_bound": low_worst,
            "step": "1e-3 relative" if mode == "sum" else "1e-3 relative (RGB), 1e-4 (alpha)",
            "resolution": "worst-case rounding of the difference <= 1e-7 max(|fd|, |g|)",
            "rays_per_cell": [min(nrays), max(nrays)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clustered2m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=float, default=None,
                    help="fraction of the rays timed on the CPU reference (default: a "
                         "quarter, or about 2.5e8 crossings if fewer)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch parameters")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ray-order", choices=("tiles", "rows"), default="tiles",
                    help="tiles: the kernels take image rays in 8x8-pixel tiles when every "
                         "view has the same row length; rows: in index order")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint measurement")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device-to-device copy of this many GiB "
                         "(reference size for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = int(args.calibrate_copy * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 8 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "rays are independent, so N GPUs would each take 1/N of the rays.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import vortrace_amd as vt
    from vortrace_amd import _lib, synthetic

    w = WORKLOADS[args.workload]
    kind, mode = w.get("kind", "image"), w["mode"]
    pos, fields, views, params = make_workload(args.workload)
    if mode == "volume":
        fields, kappa = volume_fields(pos, fields)
        log(f"[bench] kappa = {kappa:.6e}")

    t0 = time.perf_counter()
    pc = vt.ProjectionCloud(pos, fields, boundbox=[0, 100, 0, 100, 0, 100],
                            periodic=w["periodic"])
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    info = pc._cloud.tree_info()
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    ray_gen = None
    if params is not None:
        # camera rays built on the device from the views' parameters
        d_start, d_end, sizes = vt.grid.generate_projection_grids_device(params, dev)
        torch.cuda.synchronize()
        t_gen = time.perf_counter() - t0
        same = all(torch.equal(d_start[o:o + n].cpu(), torch.from_numpy(s)) and
                   torch.equal(d_end[o:o + n].cpu(), torch.from_numpy(e))
                   for (s, e), o, n in zip(views, np.cumsum([0] + sizes[:-1]), sizes))
        d_views = list(zip(torch.split(d_start, sizes), torch.split(d_end, sizes)))
        ray_gen = {"device_seconds": t_gen, "bit_identical_to_numpy": bool(same)}
        log(f"[bench] {len(sizes)} views' rays built on the device in {t_gen:.3f}s; "
            f"identical to numpy's: {same}")
        t0 = time.perf_counter()
    else:
        d_views = [(torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev))
                   for s, e in views]
        d_start, d_end = d_views[0]
    torch.cuda.synchronize()
    t_h2d = time.perf_counter() - t0
    start = np.concatenate([v[0] for v in views]) if len(views) > 1 else views[0][0]
    end = np.concatenate([v[1] for v in views]) if len(views) > 1 else views[0][1]
    R = len(start)
    n_f = 4 if mode == "volume" else 1
    n_out = 3 if mode == "volume" else 1
    log(f"[bench] tree: {info['nodes']} nodes, depth {info['max_depth']}, host build "
        f"{info['build_seconds']:.2f}s (cloud+upload {t_build:.2f}s); {len(views)} view(s), "
        f"{R} rays, H2D {t_h2d:.3f}s")

    # every view's rays are a row-major image; with one row length for all,
    # the kernels can take them in 8x8-pixel tiles (speed only)
    if params is not None:
        rows = {int(v["nres"]) if np.ndim(v["nres"]) == 0 else int(v["nres"][1])
                for v in params}
    else:
        rows = {int(w["npix"])}
    ray_row = rows.pop() if len(rows) == 1 and args.ray_order == "tiles" else 0
    pc.set_ray_rows(ray_row)
    log(f"[bench] ray order: {'8x8 tiles of %d-ray rows' % ray_row if ray_row else 'index'}")

    # Launch parameters for this workload (speed only, never a result): a
    # small sweep before the timed steps, on every ray of the step (every
    # 16th of the multi-view launch, every 4th above 2M rays).  The volume
    # kernels do not walk other lanes' intervals, so that knob is fixed there.
    tuning = None
    if not args.no_autotune:
        stride = 16 if kind == "views" else (4 if R > 2_000_000 else 1)
        helps = (128,) if mode == "volume" else (0, 128)
        tuning = vt.tuning.autotune(pc, d_start, d_end, reduction=mode, stride=stride,
                                    walk_batches=(24, 32, 48), leaf_batches=(24, 32),
                                    ray_parts=(8,), help_events=helps, repeats=1)
        tuning["stride"] = stride
        log(f"[bench] autotune on every {stride} ray(s): {tuning['best']} "
            f"({tuning['kernel_ms']:.2f} ms)")

    def forward():
        if kind == "views":  # one launch over every view's rays
            return pc.batch_projection_tensor(d_views, reduction=mode)
        return pc.projection_tensor(d_start, d_end, reduction=mode)

    out = None
    for i in range(args.warmup):
        out = forward()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = forward()
        kms.append(_lib.last_call_stats().kernel_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = _lib.last_call_stats()
    if isinstance(out, list):
        out = torch.cat(out)
    C = int(st.segments)
    ms_per_step = 1e3 * wall / args.steps
    kernel_ms = float(np.mean(kms))
    log(f"[bench] forward {args.steps} steps: {ms_per_step:.2f} ms/step, kernel "
        f"{kernel_ms:.2f} ms (min {min(kms):.2f}), {C} crossings, {st.queries} kNN queries, "
        f"{st.slow_queries} K_EXPAND, walk depth {st.max_walk_depth}, "
        f"kNN stack {st.max_knn_depth}")
    fwd = {"ms_per_step": ms_per_step, "crossings_per_s": C / (ms_per_step * 1e-3),
           "rays_per_s": R / (ms_per_step * 1e-3), "kernel_ms": kernel_ms,
           "kernel_ms_median": float(np.median(kms)), "kernel_ms_min": float(min(kms)),
           "kernel_ms_max": float(max(kms))}

    fa = None
    if kind == "fit" or (mode == "sum" and not args.no_adjoint):
        fa, grad, target = fwd_adjoint(pc, fields, d_start, d_end, args, C, R, mode)
        if kind == "fit":
            fa["fd_check"] = fd_check(pc, fields, grad, target, d_start, d_end, mode)
        del grad, target

    bytes_alg = algorithmic_bytes(R, C, len(pos), n_f, n_out)
    traffic = read_traffic(args.workload)
    traffic_fa = read_traffic(args.workload + "_fwd_adjoint")
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
    stream_gbs = stream_copy_gbs()
    log(f"[bench] device copy: {stream_gbs:.0f} GB/s (read + write)")
    if fa is not None:
        # one walk: the recording forward's bytes plus the adjoint's; beside
        # it the re-walk convention (forward twice, fields and image again)
        b_one = bytes_alg + adjoint_bytes(R, C, len(pos), n_f, n_out, sort_passes(len(pos)))
        b_rewalk = 2 * bytes_alg + C * 8 * n_f + R * 8 * n_out
        k_ms = fa["forward_kernel_ms"] + fa["adjoint_kernel_ms"]
        a_one = b_one / (k_ms * 1e-3) / 1e9
        fa["roofline"] = {
            "bound": "hbm", "kernel": "recording forward + adjoint kernels (one walk)",
            "achieved": a_one, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a_one / HBM_PEAK_GBS,
            "traffic": traffic_fa["hbm_bytes_per_launch"] if traffic_fa else None,
            "traffic_record": traffic_fa,
            "algorithmic_bytes_per_step": b_one, "kernel_ms_per_step": k_ms,
            "formula": "R(48 + 8 n_out) + C(8 n_f) + 2C(32 D + 240) [forward] + "
                       "adjoint_bytes (bench.py) [records, sort, sums]",
            "rewalk_algorithmic_bytes": b_rewalk,
            "rewalk_frac": b_rewalk / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "stream_copy_gbs": stream_gbs}
    step = fa if kind == "fit" else fwd
    if kind == "views":
        order_desc = "one launch over every view, index order"
    else:
        from vortrace_amd.autograd import _order_pays
        est = _order_pays(R, pc._cloud.lanes(pc._validate_reduction(mode)))
        order_desc = (f"8x8 tiles ({ray_row}-ray rows)" if ray_row else "index") + \
            (", runs of 64 longest first by estimated cost" if est else "")
    result = {
        "metric": METRIC,
        "value": step["crossings_per_s"],
        "unit": "cell-crossings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["desc"], "step": "forward + adjoint" if kind == "fit"
                   else "forward", "n_points": len(pos), "views": len(views), "rays": R,
                   "crossings_per_step": C, "periodic": w["periodic"],
                   "reduction": mode, "tree_depth": int(info["max_depth"]),
                   "ray_order": order_desc},
        "rays_per_s": step["rays_per_s"],
        "kernel_ms": kernel_ms,
        "kernel_crossings_per_s": C / (kernel_ms * 1e-3),
        "knn_queries_per_step": int(st.queries),
        "forward": fwd,
        "roofline": {"bound": "hbm", "kernel": "forward projection", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
                     "traffic_record": traffic,
                     "hbm_gbs_at_profile": (
                         traffic["hbm_bytes_per_launch"] / (traffic["profiled_kernel_ms"] * 1e-3)
                         / 1e9 if traffic and traffic.get("profiled_kernel_ms") else None),
                     "l2_hit_rate": traffic["l2_hit_rate"] if traffic else None,
                     "algorithmic_bytes_per_launch": bytes_alg,
                     "stream_copy_gbs": stream_gbs,
                     "frac_of_stream_copy": achieved / stream_gbs},
        "fwd_adjoint": fa,
        "tuning": tuning,
        "setup_seconds": {"tree_build_host": info["build_seconds"],
                          "cloud_and_upload": t_build, "rays_h2d": t_h2d},
        "ray_generation": ray_gen,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        frac = args.cpu_sample if args.cpu_sample else min(0.25, 2.5e8 / max(C, 1))
        gv = out.cpu().numpy().reshape(R, -1)
        result["cpu_baseline"] = cpu_baseline(pos, fields, start, end, w["periodic"],
                                              mode, frac, gv)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
