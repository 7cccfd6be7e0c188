"""Benchmark of the marching-cubes isosurface path on one MI355X.

Workload (default): 512^3-point float32 field, sum of 16 isotropic Gaussians
(centres uniform in the box, sigma ~ U[20,60] voxels, amplitude ~ U[0.5,1.5],
seed 0) plus uniform noise in [-0.01, 0.01]; isovalue 0.5; marching cubes
with duplicate-point removal, central-difference point-gradient normals and
the compact active-edge list.  One step = one forward (classify, count/scan,
generate) over the resident field into preallocated outputs.

Prints one JSON line.  `value` is Gcells/s of the forward; `roofline` is the
classify kernel (the dominant one) measured with HIP events on the stream the
kernels run on; `cpu_baseline` times the OpenMP restatement of Dax's
OpenMP-adapter pipeline on the host cores on the same field.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512, help="points per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tuning", type=str, default=None,
                    help="classify_unroll,classify_vec,scan_sub,gen_tile[,gen_xcd] (skips the autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the device copy kernel on 512 MiB (counter calibration)")
    return ap.parse_args()


def copy_probe(steps):
    """Plain copy of 512 MiB with the adapter's Copy kernel: the reference
    point for FETCH_SIZE / WRITE_SIZE scaling."""
    import torch

    from daxtoolkit_amd import _lib
    n = 512 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(steps):
        _lib.call("dax_hip_copy", dst.data_ptr(), src.data_ptr(), n, s)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "steps": steps}), flush=True)


def cpu_baseline(f_host, iso, cells, budget_s):
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    times = []
    t_start = time.perf_counter()
    T = U = 0
    while len(times) < 5 and (time.perf_counter() - t_start) < budget_s:
        sec, T, U = orc.pipeline_omp(f_host, iso, dedup=True, threads=threads)
        times.append(sec)
    med = statistics.median(times)
    return {
        "value": round(cells / med / 1e9, 4),
        "unit": "Gcells/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"full {f_host.shape[2]}x{f_host.shape[1]}x{f_host.shape[0]} field, "
                   f"OpenMP restatement of Dax's OpenMP-adapter MC pipeline (classify, scan, "
                   f"upper/lower bounds, generate, parallel sort+unique dedup, coordinates), "
                   f"median of {len(times)} runs, {med:.3f} s each, T={T} U={U}"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the volume would be split into z-slabs per GPU, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    from daxtoolkit_amd import fields, mc

    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(args.steps)
        return
    n = args.n
    dims = (n, n, n)
    iso = 0.5
    P = n ** 3
    C = (n - 1) ** 3
    params = fields.gaussian_mixture_params(0, dims)
    field = fields.gaussian_mixture_torch(dims, params)
    torch.cuda.synchronize()

    ws, cnt = mc.classify(field, iso)
    T, U = mc.read_counts(cnt)
    del ws
    buf = mc.ForwardBuffers(n, n, n, U, T, normals=True, edges=True)
    if args.tuning:
        tuned = [int(x) for x in args.tuning.split(",")]
        mc.set_tuning((n, n, n), tuned)
        tuned_ms = [0.0, 0.0, 0.0]
    else:
        tuned, tuned_ms = mc.autotune(field, iso, buf, reps=10)

    for _ in range(args.warmup):
        mc.forward_into(field, iso, buf)
    torch.cuda.synchronize()
    mc.read_counts(buf.counts, buf.cap_points, buf.cap_tris)

    # timed region: K forwards
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        mc.forward_into(field, iso, buf)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = e0.elapsed_time(e1) / args.steps
    T2, U2 = mc.read_counts(buf.counts, buf.cap_points, buf.cap_tris)
    assert (T2, U2) == (T, U)

    # per-kernel durations (same work, one kernel per call, events between)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    for k in range(args.steps):
        mc.forward_phases(field, iso, buf, events=ev[k])
    torch.cuda.synchronize()
    k_ms = [statistics.mean(e[i].elapsed_time(e[i + 1]) for e in ev) for i in range(3)]

    b_classify = 4 * P + P // 8                      # field read once + inside-bits written
    b_fwd = 4 * P + 12 * U + 12 * U + 4 * 3 * T + 4 * U  # field + verts + normals + conn + edges
    achieved = b_classify / (k_ms[0] * 1e-3) / 1e9
    fwd_gbs = b_fwd / (ms * 1e-3) / 1e9

    line = {
        "metric": "MC Gcells/s at 512^3 fwd and fwd+bwd; % HBM roofline; batched volumes/s",
        "value": round(C / (ms * 1e-3) / 1e9, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"marching cubes forward, {n}^3 points, 16-Gaussian mixture "
                               f"(seed 0) + U(-0.01,0.01) noise, iso 0.5, dedup, "
                               f"central-difference normals, compact edge list",
                   "points": P, "cells": C, "triangles": T, "vertices": U},
        "tris_per_s": round(T / (ms * 1e-3), 1),
        "fwd_bytes": b_fwd,
        "fwd_GBps": round(fwd_gbs, 1),
        "fwd_frac_hbm_peak": round(fwd_gbs / HBM_PEAK_GBS, 4),
        "kernel_ms": {"classify": round(k_ms[0], 5), "count_scan": round(k_ms[1], 5),
                      "generate": round(k_ms[2], 5)},
        "wall_s": round(wall, 4),
        "tuning": {"classify_unroll": tuned[0], "classify_vec": tuned[1], "scan_sub": tuned[2],
                   "gen_tile": tuned[3], "gen_xcd": (tuned + [1])[4],
                   "autotune_ms": [round(x, 5) for x in tuned_ms]},
        "roofline": {"bound": "hbm", "kernel": "k_classify", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
    }
    if not args.no_cpu_baseline:
        f_host = field.cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(f_host, iso, C, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
