"""Benchmark: cell-crossings/s and rays/s of the forward projection on one GPU.

Default workload (one step = one full image):
  2,000,000-point clustered periodic mesh (box 100, seed 2), 512^2 rays along
  +z over [0, 100], reduction 'sum', FP64 throughout.
Inputs (tree, fields, rays) are resident on the device before the timed
region; tree build and host->device copies are reported separately.

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
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    from vortrace_amd import grid, synthetic
    w = WORKLOADS[name]
    t0 = time.perf_counter()
    if name == "uniform100k":
        pos, fields = synthetic.uniform_mesh(w["n"], seed=1)
    else:
        pos, fields = synthetic.clustered_mesh(w["n"], seed=2)
    log(f"[bench] mesh {name}: {len(pos)} points in {time.perf_counter() - t0:.1f}s "
        f"(sha {synthetic.digest(pos, fields)})")
    s, e = grid.generate_projection_grid([0.0, 100.0], w["npix"], [0.0, 100.0], None)
    return pos, fields, s.reshape(-1, 3), e.reshape(-1, 3)


def algorithmic_bytes(R, C, n_points, n_f, n_out):
    """R(48 + 8 n_out) + C(8 n_f) + 2C(32 D + 240), D = ceil(log2(N/10)):
    ray endpoints and outputs, one field gather per crossing, and two kNN
    queries per crossing, each one root-to-leaf descent of 32-byte nodes plus
    one 10-point leaf of 24-byte generators."""
    D = math.ceil(math.log2(n_points / 10.0))
    return R * (48 + 8 * n_out) + C * (8 * n_f) + 2 * C * (32 * D + 240)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE runs, scaled by a copy-kernel calibration) and the L2 hit rate
    (TCC_HIT / (TCC_HIT + TCC_MISS)) of the same kernel, if recorded."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch"), d.get("l2_hit_rate")


def cpu_baseline(pos, fields, start, end, periodic, mode, frac, gpu_vals):
    """The reference's C++ core (oracle/_ref) on a strided sample of the rays."""
    from oracle import ref
    if not ref.available():
        return None
    stride = max(1, int(round(1.0 / frac)))
    idx = np.arange(0, len(start), stride)
    c = ref.RefCloud(pos, fields, [0, 100, 0, 100, 0, 100], periodic=periodic)
    c.project(start[idx[:64]], end[idx[:64]], mode)  # warm
    t0 = time.perf_counter()
    vals, ids, _, _, off = c.detailed(start[idx], end[idx], mode)
    t_det = time.perf_counter() - t0
    t0 = time.perf_counter()
    vals = c.project(start[idx], end[idx], mode)
    dt = time.perf_counter() - t0
    crossings = int(off[-1])
    same = bool(np.array_equal(vals.reshape(gpu_vals[idx].shape), gpu_vals[idx]))
    close = bool(np.allclose(vals.reshape(gpu_vals[idx].shape), gpu_vals[idx],
                             rtol=1e-12, atol=0))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"[bench] cpu reference: {len(idx)} rays, {crossings} crossings in "
        f"{dt:.2f}s ({threads} threads); values bit-identical: {same}")
    return {"value": crossings / dt, "unit": "cell-crossings/s", "cores": threads,
            "kind": "reference",
            "sample": f"every {stride}th ray of the same image ({len(idx)} rays, "
                      f"{crossings} crossings), reference C++ core "
                      f"(makeProjection, OpenMP dynamic,256)",
            "rays_per_s": len(idx) / dt,
            "values_bit_identical_to_gpu": same,
            "values_within_1e-12": close}


def fwd_adjoint(pc, rho, d_start, d_end, args, C, R):
    """Forward plus adjoint on the same rays: one step evaluates
    L = sum over pixels of (proj(rho) - target)^2 and dL/drho through
    torch.autograd (projection kernel forward, adjoint kernel backward), with
    target = proj(rho * exp(0.1 N(0,1) per cell))."""
    import torch
    import vortrace_amd as vt
    from vortrace_amd import synthetic
    dev = d_start.device
    rng = np.random.default_rng(55)
    noisy = rho * np.exp(0.1 * synthetic.normal(rng, len(rho)))
    pc.set_fields(noisy)
    target = pc.projection_tensor(d_start, d_end)[:, 0].clone()
    rho_t = torch.from_numpy(np.ascontiguousarray(rho, dtype=np.float64)).to(dev)

    split = {"forward_ms": [], "backward_ms": []}

    def step():
        f = rho_t.clone().requires_grad_(True)
        t0 = time.perf_counter()
        p = vt.autograd.project(pc, f, d_start, d_end, "sum")
        loss = ((p - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        split["forward_ms"].append(1e3 * (t1 - t0))
        split["backward_ms"].append(1e3 * (time.perf_counter() - t1))
        return f.grad, pc._adjoint_stats

    for i in range(args.warmup):
        step()
        log(f"[bench] fwd+adjoint warmup {i + 1}/{args.warmup} done")
    split["forward_ms"].clear()
    split["backward_ms"].clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    for i in range(args.steps):
        g, st = step()
        kms.append(st.kernel_ms)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    nz = int(torch.count_nonzero(g).item())
    fwd_ms, bwd_ms = float(np.mean(split["forward_ms"])), float(np.mean(split["backward_ms"]))
    log(f"[bench] fwd+adjoint {args.steps} steps: {ms:.2f} ms/step (forward {fwd_ms:.2f}, "
        f"backward {bwd_ms:.2f}), adjoint kernels {float(np.mean(kms)):.2f} ms, {nz} cells "
        f"with a gradient")
    pc.set_fields(rho)
    return {"ms_per_step": ms, "crossings_per_s": C / (ms * 1e-3), "rays_per_s": R / (ms * 1e-3),
            "adjoint_kernel_ms": float(np.mean(kms)), "forward_ms": fwd_ms,
            "backward_ms": bwd_ms, "cells_with_gradient": nz,
            "step": "L = sum (proj(rho) - target)^2 and dL/drho by torch.autograd "
                    "(forward kernel recording its segments + adjoint kernels over the "
                    "records), FP64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clustered2m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=float, default=0.25,
                    help="fraction of the rays timed on the CPU reference")
    ap.add_argument("--no-cpu", action="store_true")
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
    pos, fields, start, end = make_workload(args.workload)
    if w["mode"] == "volume":
        # kappa: median ray opacity 0.5 on a 64^2 probe of the same view
        pc_probe = vt.ProjectionCloud(pos, fields, boundbox=[0, 100, 0, 100, 0, 100],
                                      periodic=True)
        ps, pe = vt.grid.generate_projection_grid([0.0, 100.0], 64, [0.0, 100.0], None)
        _, ids, se, sx, off = pc_probe.segments_tensor(
            torch.from_numpy(ps.reshape(-1, 3)), torch.from_numpy(pe.reshape(-1, 3)))
        kappa = synthetic.kappa_for_median_opacity(
            fields, ids.cpu().numpy(), (sx - se).cpu().numpy(), off.cpu().numpy())
        fields = synthetic.rgba_transfer(fields, kappa)
        del pc_probe
        log(f"[bench] kappa = {kappa:.6e}")

    t0 = time.perf_counter()
    pc = vt.ProjectionCloud(pos, fields, boundbox=[0, 100, 0, 100, 0, 100],
                            periodic=w["periodic"])
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    info = pc._cloud.tree_info()
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    d_start = torch.from_numpy(start).to(dev)
    d_end = torch.from_numpy(end).to(dev)
    torch.cuda.synchronize()
    t_h2d = time.perf_counter() - t0
    R = len(start)
    n_f = 4 if w["mode"] == "volume" else 1
    n_out = 3 if w["mode"] == "volume" else 1
    log(f"[bench] tree: {info['nodes']} nodes, depth {info['max_depth']}, host build "
        f"{info['build_seconds']:.2f}s (cloud+upload {t_build:.2f}s); rays H2D {t_h2d:.3f}s")

    out = None
    for i in range(args.warmup):
        out = pc.projection_tensor(d_start, d_end, reduction=w["mode"])
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = pc.projection_tensor(d_start, d_end, reduction=w["mode"])
        kms.append(_lib.last_call_stats().kernel_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = _lib.last_call_stats()
    C = int(st.segments)
    ms_per_step = 1e3 * wall / args.steps
    kernel_ms = float(np.mean(kms))
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, kernel {kernel_ms:.2f} ms "
        f"(min {min(kms):.2f}), {C} crossings, {st.queries} kNN queries, "
        f"{st.slow_queries} K_EXPAND, walk depth {st.max_walk_depth}, "
        f"kNN stack {st.max_knn_depth}")

    fa = None
    if w["mode"] == "sum" and not args.no_adjoint:
        fa = fwd_adjoint(pc, fields, d_start, d_end, args, C, R)

    bytes_alg = algorithmic_bytes(R, C, len(pos), n_f, n_out)
    traffic, l2_hit = read_traffic(args.workload)
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": C / (ms_per_step * 1e-3),
        "unit": "cell-crossings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["desc"], "n_points": len(pos), "rays": R,
                   "crossings_per_step": C, "periodic": w["periodic"],
                   "reduction": w["mode"], "tree_depth": int(info["max_depth"])},
        "rays_per_s": R / (ms_per_step * 1e-3),
        "kernel_ms": kernel_ms,
        "kernel_crossings_per_s": C / (kernel_ms * 1e-3),
        "knn_queries_per_step": int(st.queries),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "hbm_gbs_measured": (traffic / (kernel_ms * 1e-3) / 1e9
                                          if traffic else None),
                     "l2_hit_rate": l2_hit,
                     "algorithmic_bytes_per_launch": bytes_alg},
        "fwd_adjoint": fa,
        "setup_seconds": {"tree_build_host": info["build_seconds"],
                          "cloud_and_upload": t_build, "rays_h2d": t_h2d},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        gv = out.cpu().numpy().reshape(R, -1)
        result["cpu_baseline"] = cpu_baseline(pos, fields, start, end, w["periodic"],
                                              w["mode"], args.cpu_sample, gv)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
