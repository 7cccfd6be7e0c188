"""Benchmark of the marching-cubes isosurface path on one MI355X.

Headline workload: 512^3-point float32 field, sum of 16 isotropic Gaussians
(centres uniform in the box, sigma ~ U[20,60] voxels, amplitude ~ U[0.5,1.5],
seed 0) plus uniform noise in [-0.01, 0.01]; isovalue 0.5; marching cubes
with duplicate-point removal, central-difference point-gradient normals and
the compact active-edge list.  One step = one forward (classify, count/scan,
generate) over the resident field into preallocated outputs.

Timed regions are the K steps captured as one HIP graph and launched once
(bracketed by device synchronisation), so host-side jitter between Python
calls cannot leave the GPU idle; the same K steps launched one by one from
Python are reported beside it.

Prints one JSON line.  `value` is Gcells/s of the forward; `roofline` is
the forward graph (its algorithmic bytes over its HIP-event time on the
stream the kernels run on), with each kernel's own time, bytes and measured
HBM traffic under `kernels`; `cpu_baseline` times the OpenMP restatement of
Dax's OpenMP-adapter pipeline on the host cores on the same field.  Beside
the headline the line carries
  fwd_bwd  the same field, forward + adjoint (upstream dL/dvertex and
           dL/dnormal N(0,1), seed 1), and its ratio to the forward;
  fwd_bwd_256  the differentiable workload: the same at 256^3, hot (steps back
           to back, field in the MALL) and cold (MALL flushed before each
           step; the headline ratio);
  slice_512  a slice of the 512^3 grid by a tilted plane (no field read),
           with the OpenMP restatement beside it;
  tetrahedra_128  marching tetrahedra over the 12.3 M Freudenthal
           tetrahedra of a 128^3 grid, with the serial restatement beside it;
  hexahedra_256  marching cubes on the unstructured hexahedral grid of the
           256^3 grid's voxels (the general dedup path), with the serial
           restatement beside it;
  threshold_512  the reference's Threshold benchmark pipeline at 512^3
           (|p| field, [0, 100], duplicate removal, CompactPointField), with
           the serial restatement beside it;
  vtk_write_512  the 512^3 contour written as a VTK legacy file by the
           library's host writer, with the reference's ofstream writer
           (restated) beside it;
  batched  1024 volumes with per-axis extents U{32..160}, 4-16 Gaussians each,
           iso U[0.3,0.7]: the batched forward and forward+backward replayed
           as captured graphs, against a native launch loop of single-volume
           forwards (+ backwards) into the same outputs;
  large    1024^3 Mandelbulb (power 8, 10 iterations, bailout 2, iso 0.5)
           forward and forward+backward (--no-large skips it);
  dispatcher_pipeline_512  the reference benchmark's timed region
           (Benchmarks/MarchingCubes/Pipeline.h:110-132: classify, then
           generate with duplicate removal, |p| field at 512^3, iso 100)
           through the C++ dispatchers, as a Dax user calls them (the
           program tests/cpp/bin/BenchmarkPipelineMarchingCubesHip).
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


def _kernel_variants(n, tuned):
    """Template names of the three forward kernels a tuning launches."""
    unroll, vec, sub, tile = tuned[0], tuned[1], tuned[2], tuned[3]
    cls = (f"k_classify_v4<{unroll}>" if vec == 4 else
           f"k_classify<{unroll}, {'true' if n % 64 == 0 else 'false'}>")
    return cls, f"k_count<{sub}>", f"k_generate<{tile}, FieldArray>"


def kernel_sources_sha256():
    """SHA-256 over the library's kernel sources (daxtoolkit_amd/csrc/*.hip
    and *.h, by name), the build a traffic summary was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted((ROOT / "daxtoolkit_amd" / "csrc").glob("*")):
        if f.suffix in (".hip", ".h"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()


TRAFFIC_NOTE = None


def _measured_traffic(n, tuned):
    """HBM bytes per dispatch of the forward's kernels (by role) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the 512^3
    workload, scaled by the copy-kernel calibration; None for another
    workload, a kernel variant the summary does not hold, or a summary
    measured on other kernel sources (stale)."""
    global TRAFFIC_NOTE
    path = ROOT / "profiles" / "forward_512_traffic.json"
    if n != 512 or not path.exists():
        return None
    j = json.loads(path.read_text())
    if j.get("sources_sha256") != kernel_sources_sha256():
        TRAFFIC_NOTE = ("profiles/forward_512_traffic.json was measured on other kernel "
                        "sources: traffic not reported")
        return None
    TRAFFIC_NOTE = "profiles/forward_512_traffic.json, measured on these kernel sources"
    d = j["hbm_bytes_per_dispatch"]
    roles = ("k_classify_v4", "k_count", "k_generate")
    return {r: d.get(v) for r, v in zip(roles, _kernel_variants(n, tuned))}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512, help="points per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tuning", type=str, default=None,
                    help="classify_unroll,classify_vec,scan_sub,gen_tile[,gen_xcd] "
                         "(skips the autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the device copy kernel on 512 MiB (counter calibration)")
    ap.add_argument("--batch", type=int, default=1024, help="volumes of the batched run (0: skip)")
    ap.add_argument("--no-bwd", action="store_true", help="skip the forward+backward timing")
    ap.add_argument("--no-large", action="store_true", help="skip the 1024^3 Mandelbulb")
    ap.add_argument("--only", choices=["fwd", "batch", "large", "diff", "slice", "tets", "hex",
                                       "threshold", "vtk", "dispatch"],
                    default=None,
                    help="run one section only (profiling); prints its JSON")
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


def _omp_baseline(run, units, unit, what, budget_s):
    """Median of up to 5 runs (within budget_s) of an OpenMP restatement:
    run(threads) -> (seconds, *counts).  Returns the cpu_baseline entry and
    the counts of the last run."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    times, counts = [], ()
    t_start = time.perf_counter()
    while len(times) < 5 and (len(times) == 0 or time.perf_counter() - t_start < budget_s):
        sec, *counts = run(threads)
        times.append(sec)
    med = statistics.median(times)
    return {"value": round(units / med / 1e9, 4), "unit": unit, "cores": threads, "kind": "port",
            "sample": f"the full input, {what}, median of {len(times)} runs, {med:.3f} s each"}, \
        tuple(counts)


def _timed(fn, steps, per_step=None):
    """Mean ms of fn() over `steps` calls, HIP events on the current stream
    (one event between consecutive steps; the per-step times go into
    `per_step` when a list is given)."""
    import torch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    ev[0].record()
    for i in range(steps):
        fn()
        ev[i + 1].record()
    torch.cuda.synchronize()
    if per_step is not None:
        per_step.extend(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
    return ev[0].elapsed_time(ev[steps]) / steps


def _graph_ms(fn, steps, stream):
    """Mean ms of fn() when `steps` calls are captured as one graph."""
    import torch

    from daxtoolkit_amd import mc
    g = mc.capture(lambda: [fn() for _ in range(steps)], stream)
    try:
        with torch.cuda.stream(stream):
            mc.launch_graph(g, stream)
            return _timed(lambda: mc.launch_graph(g, stream), 1) / steps
    finally:
        mc.destroy_graph(g)


def _progress(msg):
    print(f"[bench] {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def _upstream(U, seed):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    dv = torch.randn((max(U, 1), 3), generator=gen, device="cuda", dtype=torch.float32)
    dn = torch.randn((max(U, 1), 3), generator=gen, device="cuda", dtype=torch.float32)
    return dv, dn


def _flusher():
    """A launch that moves 512 MiB through the HBM (a 256 MiB copy): between
    timed steps it leaves nothing of the previous step in the 256 MB MALL."""
    import torch

    from daxtoolkit_amd import _lib
    n = 256 << 20
    src = torch.zeros(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    return lambda: _lib.call("dax_hip_copy", dst.data_ptr(), src.data_ptr(), n,
                             torch.cuda.current_stream().cuda_stream)


def _cold_ms(step, flush, steps, stream):
    """Mean ms of step() with the MALL flushed before each: the graph of
    (flush, step) x K minus the graph of flush x K."""
    return max(0.0, _graph_ms(lambda: (flush(), step()), steps, stream) -
               _graph_ms(flush, steps, stream))


def single_volume(field, iso, label, args, tuning=None, cold=False):
    """Forward (+ phases, + fwd+bwd) of one resident volume; with cold, also
    the forward and forward+backward with the MALL flushed before each step."""
    import torch

    from daxtoolkit_amd import mc
    nz, ny, nx = field.shape
    P, C = nx * ny * nz, (nx - 1) * (ny - 1) * (nz - 1)
    ws, cnt = mc.classify(field, iso)
    T, U = mc.read_counts(cnt)
    del ws
    buf = mc.ForwardBuffers(nx, ny, nz, max(U, 1), max(T, 1), normals=True, edges=True)
    _progress(f"{label}: classified; tuning")
    if tuning:
        mc.set_tuning((nx, ny, nz), tuning)
        tuned, tuned_ms = mc.get_tuning((nx, ny, nz)), [0.0, 0.0, 0.0]
    else:
        tuned, tuned_ms = mc.autotune(field, iso, buf, reps=10)
    for _ in range(args.warmup):
        mc.forward_into(field, iso, buf)
    torch.cuda.synchronize()
    print(f"[bench] {label}: T={T} U={U} tuning={tuned}", file=sys.stderr, flush=True)
    # timed region: the K forwards captured as one graph (one host launch, so
    # host-side jitter cannot idle the GPU between steps)
    s = torch.cuda.Stream()
    g_fwd = mc.capture(lambda: [mc.forward_into(field, iso, buf) for _ in range(args.steps)], s)
    with torch.cuda.stream(s):
        mc.launch_graph(g_fwd, s)
        wall0 = time.perf_counter()
        ms = _timed(lambda: mc.launch_graph(g_fwd, s), 1) / args.steps
        wall = time.perf_counter() - wall0
    mc.destroy_graph(g_fwd)
    _progress(f"{label}: forward graph {ms:.4f} ms per step")
    steps_ms = []  # the same K forwards launched one by one from Python
    loop_ms = _timed(lambda: mc.forward_into(field, iso, buf), args.steps, steps_ms)
    assert mc.read_counts(buf.counts, buf.cap_points, buf.cap_tris) == (T, U)
    # per-kernel durations, each from K launches captured as one graph:
    # classify alone; classify + count/scan (count/scan needs the state
    # classify clears) minus classify; generate alone (it only reads the
    # workspace)
    t_cls = _graph_ms(lambda: mc.classify_into(field, iso, buf), args.steps, s)
    t_cs = _graph_ms(lambda: (mc.classify_into(field, iso, buf), mc.count_scan_into(buf)),
                     args.steps, s) - t_cls
    t_gen = _graph_ms(lambda: mc.generate_into(field, iso, buf), args.steps, s)
    k_ms = [t_cls, t_cs, t_gen]
    _progress(f"{label}: kernels " + ", ".join(f"{x:.4f}" for x in k_ms) + " ms")
    r = {"P": P, "C": C, "T": T, "U": U, "ms": ms, "wall": wall, "k_ms": k_ms,
         "tuned": tuned, "tuned_ms": tuned_ms, "steps_ms": steps_ms, "loop_ms": loop_ms}
    if not args.no_bwd:
        dv, dn = _upstream(U, 1)
        scratch = torch.empty(max(int(mc._lib.load().dax_hip_mc_backward_scratch_bytes(U)), 1),
                              dtype=torch.uint8, device="cuda")
        dldf = torch.empty_like(field)

        def fb():
            mc.forward_into(field, iso, buf)
            mc.backward_into(field, iso, buf, dv, dn, scratch, dldf)
        for _ in range(2):
            fb()
        torch.cuda.synchronize()
        _progress(f"{label}: eager forward+backward done")
        g_fb = mc.capture(lambda: [fb() for _ in range(args.steps)], s)
        g_b = mc.capture(lambda: [mc.backward_into(field, iso, buf, dv, dn, scratch, dldf)
                                  for _ in range(args.steps)], s)
        with torch.cuda.stream(s):
            mc.launch_graph(g_fb, s)
            r["fb_ms"] = _timed(lambda: mc.launch_graph(g_fb, s), 1) / args.steps
            r["bwd_ms"] = _timed(lambda: mc.launch_graph(g_b, s), 1) / args.steps
        mc.destroy_graph(g_fb)
        mc.destroy_graph(g_b)
        print(f"[bench] {label}: fwd {ms:.4f} ms, fwd+bwd {r['fb_ms']:.4f} ms",
              file=sys.stderr, flush=True)
    if cold:
        flush = _flusher()
        r["ms_cold"] = _cold_ms(lambda: mc.forward_into(field, iso, buf), flush, args.steps, s)
        if not args.no_bwd:
            r["fb_ms_cold"] = _cold_ms(fb, flush, args.steps, s)
        _progress(f"{label}: cold (MALL flushed) fwd {r['ms_cold']:.4f} ms"
                  + (f", fwd+bwd {r['fb_ms_cold']:.4f} ms" if "fb_ms_cold" in r else ""))
    return r


def batched(args):
    """1024 uneven volumes: batched graphs vs a native per-volume loop."""
    import torch

    from daxtoolkit_amd import fields, mc
    vb = fields.volume_batch(args.batch, seed=0)
    flat = fields.volume_batch_torch(vb)
    b = mc.Batch(vb.dims, vb.isos)
    T, U = b.sizes(flat)
    P = int(flat.numel())
    C = int(sum((d[0] - 1) * (d[1] - 1) * (d[2] - 1) for d in vb.dims))
    print(f"[bench] batch: {args.batch} volumes, {P} points, T={T} U={U}", file=sys.stderr,
          flush=True)
    bufs = b.alloc(U, T)
    tuned, tuned_ms = b.autotune(flat, bufs, reps=3)
    _progress(f"batch: autotuned {tuned} (classify, count/scan, generate "
              + ", ".join(f"{x:.3f}" for x in tuned_ms) + " ms)")
    dv, dn = _upstream(U, 2)
    scratch = torch.empty(max(b.scratch_bytes(U), 1), dtype=torch.uint8, device="cuda")
    dldf = torch.empty_like(flat)
    b.forward_into(flat, bufs)
    b.backward(flat, bufs, dv, dn, scratch, dldf)
    torch.cuda.synchronize()
    offs = b.vol_offsets.cpu().numpy()
    s = torch.cuda.Stream()
    g_f = mc.capture(lambda: b.forward_into(flat, bufs), s)
    g_fb = mc.capture(lambda: (b.forward_into(flat, bufs),
                               b.backward(flat, bufs, dv, dn, scratch, dldf)), s)
    steps = max(3, min(args.steps, 10))
    try:
        with torch.cuda.stream(s):
            for g in (g_f, g_fb):
                mc.Batch.launch(g, s)
            ms_f = _timed(lambda: mc.Batch.launch(g_f, s), steps)
            ms_fb = _timed(lambda: mc.Batch.launch(g_fb, s), steps)
    finally:
        mc.Batch.destroy(g_f)
        mc.Batch.destroy(g_fb)
    assert mc.read_counts(b.counts) == (T, U)
    lbufs = b.alloc(U, T)
    ldldf = torch.empty_like(flat)
    lws = b.loop_workspace()
    lcnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    b.loop(flat, lbufs, offs, lws, lcnt)
    ms_lf = _timed(lambda: b.loop(flat, lbufs, offs, lws, lcnt), steps)
    grads = (dv, dn, torch.empty_like(scratch), ldldf)
    b.loop(flat, lbufs, offs, lws, lcnt, grads=grads)
    ms_lfb = _timed(lambda: b.loop(flat, lbufs, offs, lws, lcnt, grads=grads), steps)
    same = all(torch.equal(bufs[k][:n], lbufs[k][:n]) for k, n in
               (("vertices", U), ("normals", U), ("conn", T), ("edges", U)))
    same = bool(same and torch.equal(dldf, ldldf))
    return {
        "workload": f"{args.batch} volumes, per-axis point extents U{{32..160}} (seed 0), "
                    f"4-16 Gaussians each, iso U[0.3,0.7], dedup, normals, edge list",
        "volumes": args.batch, "points": P, "cells": C, "triangles": T, "vertices": U,
        "graph_fwd_ms": round(ms_f, 4), "graph_fwd_bwd_ms": round(ms_fb, 4),
        "loop_fwd_ms": round(ms_lf, 4), "loop_fwd_bwd_ms": round(ms_lfb, 4),
        "speedup_fwd": round(ms_lf / ms_f, 3), "speedup_fwd_bwd": round(ms_lfb / ms_fb, 3),
        "volumes_per_s_fwd": round(args.batch / (ms_f * 1e-3), 1),
        "volumes_per_s_fwd_bwd": round(args.batch / (ms_fb * 1e-3), 1),
        "Gcells_per_s_fwd": round(C / (ms_f * 1e-3) / 1e9, 3),
        "loop_outputs_bit_identical": same, "steps": steps,
        "tuning": {"classify_unroll": tuned[0], "classify_blocks_per_cu": tuned[1],
                   "scan_threads": tuned[2], "gen_blocks_per_cu": tuned[3], "gen_xcd": tuned[4],
                   "autotune_phase_ms": [round(x, 4) for x in tuned_ms]},
    }


def differentiable_256(args):
    """The differentiable workload: 256^3, the same 16-Gaussian field (seed
    0), forward and forward+backward (upstream N(0,1))."""
    import torch

    from daxtoolkit_amd import fields
    n = 256
    field = fields.gaussian_mixture_torch((n, n, n), fields.gaussian_mixture_params(0, (n, n, n)))
    torch.cuda.synchronize()
    r = single_volume(field, 0.5, "256^3", args, cold=True)
    return {"workload": "differentiable marching cubes, 256^3 points, 16-Gaussian mixture (seed 0) "
                        "+ U(-0.01,0.01) noise, iso 0.5, upstream dL/dvertex, dL/dnormal N(0,1)",
            "triangles": r["T"], "vertices": r["U"],
            "hot": {"note": "K steps back to back: the 64 MB field stays in the 256 MB MALL",
                    "fwd_ms": round(r["ms"], 5), "fwd_bwd_ms": round(r["fb_ms"], 5),
                    "bwd_ms": round(r["bwd_ms"], 5),
                    "fwd_bwd_over_fwd": round(r["fb_ms"] / r["ms"], 3)},
            "cold": {"note": "512 MiB of copy traffic before each step (not timed): every step "
                             "reads the field from HBM",
                     "fwd_ms": round(r["ms_cold"], 5), "fwd_bwd_ms": round(r["fb_ms_cold"], 5),
                     "fwd_bwd_over_fwd": round(r["fb_ms_cold"] / r["ms_cold"], 3)},
            "fwd_ms": round(r["ms_cold"], 5), "fwd_bwd_ms": round(r["fb_ms_cold"], 5),
            "fwd_bwd_over_fwd": round(r["fb_ms_cold"] / r["ms_cold"], 3), "tuning": r["tuned"]}


def slice_512(args):
    """A slice of the 512^3 uniform grid by a tilted plane (SliceCount /
    SliceGenerate semantics; the plane distance is computed in the kernels,
    no field is read), graph-timed like the forward; beside it the OpenMP
    restatement of the MC pipeline on the materialised plane values."""
    import torch

    from daxtoolkit_amd import mc
    n = 512
    dims = (n, n, n)
    po, pn = (128.0, 128.0, 100.0), (0.36, 0.48, 0.8)
    ws_cnt = mc.slice(dims, po, pn, normals=False)
    T, U = ws_cnt.num_triangles, ws_cnt.num_points
    del ws_cnt
    buf = mc.ForwardBuffers(n, n, n, max(U, 1), max(T, 1), normals=False, edges=True)
    for _ in range(args.warmup):
        mc.slice_into(dims, po, pn, buf)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    ms = _graph_ms(lambda: mc.slice_into(dims, po, pn, buf), args.steps, s)
    assert mc.read_counts(buf.counts, buf.cap_points, buf.cap_tris) == (T, U)
    C = (n - 1) ** 3
    out = {"workload": f"slice of the {n}^3 grid (origin 0, spacing 1) by the plane "
                       f"dot({pn}, c - {po}) = dot(n, o) (SliceCount/SliceGenerate), dedup, "
                       f"edge list, no normals",
           "triangles": T, "vertices": U, "ms": round(ms, 5),
           "Gcells_per_s": round(C / (ms * 1e-3) / 1e9, 3),
           "tris_per_s": round(T / (ms * 1e-3), 1),
           "out_bytes": 12 * U + 12 * T + 4 * U + n ** 3 // 8}
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        f, iso = orc.slice_field(dims, po, pn)
        cb = cpu_baseline(f, float(iso), C, args.cpu_seconds)
        cb["sample"] += " (the plane values materialised beforehand, not timed)"
        out["cpu_baseline"] = cb
    _progress(f"slice {n}^3: {ms:.4f} ms, T={T}")
    return out


def tetrahedra_128(args):
    """Marching tetrahedra over the Freudenthal tetrahedra of a 128^3 grid
    (6 per voxel, as the reference's test grids) on the 128^3 sphere-distance
    field, iso 40, dedup on; the call is synchronous (counts and the unique
    size come back to the host), so each rep is timed whole with HIP events.
    Beside it the oracle's serial restatement on the same input."""
    import torch

    from daxtoolkit_amd import fields, mc
    n = 128
    tets, pts = fields.voxel_cells_torch((n, n, n), "tetrahedron")
    f_h = fields.sphere_distance(n)
    iso = 40.0
    f = torch.from_numpy(f_h.reshape(-1)).cuda()
    out = mc.tetrahedra(tets, f, pts, iso)
    T, U = out.num_triangles, out.num_points
    reps = max(3, min(args.steps, 10))
    ms = _timed(lambda: mc.tetrahedra(tets, f, pts, iso), reps)
    C = int(tets.shape[0])
    res = {"workload": f"marching tetrahedra, Freudenthal tetrahedra of a {n}^3 grid "
                       f"({C} tetrahedra), sphere distance |p - (63.5,63.5,63.5)|, iso 40, "
                       f"dedup",
           "tetrahedra": C, "triangles": T, "vertices": U, "ms": round(ms, 4),
           "Gtets_per_s": round(C / (ms * 1e-3) / 1e9, 3),
           "alg_bytes": _cells_alg_bytes(C, 4, T, U)}
    res["GBps"] = round(res["alg_bytes"] / (ms * 1e-3) / 1e9, 1)
    res["frac_hbm_peak"] = round(res["GBps"] / HBM_PEAK_GBS, 4)
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        tets_h, pts_h = tets.cpu().numpy(), pts.cpu().numpy()
        cb, counts = _omp_baseline(
            lambda th: orc.unstructured_pipeline_omp(tets_h, f_h.reshape(-1), pts_h, iso,
                                                     dedup=True, threads=th),
            C, "Gtets/s", "OpenMP restatement of Dax's OpenMP-adapter marching-tetrahedra "
            "sequence (classify, scan, upper/lower bounds, interpolation records, parallel "
            "sort+unique dedup, coordinates)", args.cpu_seconds)
        assert counts == (T, U)
        res["cpu_baseline"] = cb
    _progress(f"tetrahedra {n}^3: {ms:.3f} ms, T={T}")
    return res


def _cells_alg_bytes(C, nv, T, U):
    """Algorithmic HBM bytes of one unstructured-grid contour: the count pass
    reads the connectivity (4*nv B per cell) and writes a count (4 B); the
    scan reads the counts and writes the offsets (8 B per cell); generate
    (a thread per triangle) reads its cell's ids and offset (4*nv + 4 B per
    triangle) and writes an 8-byte edge key per corner; dedup sorts and looks
    up the 3T keys (counted once: read + write, 16 B) and writes U points
    (first, second, weight, xyz: 24 B).  Field and coordinate gathers stay
    L2/MALL resident at these sizes and are left out, as are the binary
    searches' other offset reads."""
    return C * (4 * nv + 4 + 8) + T * (4 * nv + 4) + 3 * T * (8 + 16) + U * 24


def threshold_512(args):
    """The reference's Threshold benchmark pipeline (Benchmarks/Threshold/
    Pipeline.h:163-190, main.cxx: origin 0, spacing 1) at 512^3: the field
    is the point magnitude |p| (made beforehand, as the reference does with
    DispatcherMapField<Magnitude> before its timer starts); timed:
    ThresholdCount [0, 100], DispatcherGenerateTopology with duplicate
    removal, and CompactPointField of the field.  The reference's int32
    point mask is internal to that pipeline and is not kept here.  Synchronous like the
    reference (the cell and point counts come back to the host)."""
    import torch

    from daxtoolkit_amd import topology
    n = 512
    ax = torch.arange(n, device="cuda", dtype=torch.float32)
    z, y, x = torch.meshgrid(ax, ax, ax, indexing="ij")
    f = torch.sqrt(x * x + y * y + z * z).contiguous()
    del x, y, z
    lo, hi = 0.0, 100.0

    def run():
        out = topology.threshold(f, lo, hi, point_mask=False)  # the mask is internal there
        return out, out.compact_point_field(f.reshape(-1))

    out, vals = run()
    T, U = out.num_cells, out.num_points
    reps = max(3, min(args.steps, 10))
    ms = _timed(run, reps)
    P, C = n ** 3, (n - 1) ** 3
    # the bytes the pipeline cannot avoid: the field read, the counts written
    # (ThresholdCount's output), the field read again by CompactPointField,
    # and the outputs (8 ids per cell, 3 coordinates and 1 value per point)
    alg = 4 * P + 4 * C + 4 * P + 32 * T + 16 * U
    res = {"workload": f"threshold of the {n}^3 grid (origin 0, spacing 1), field |p|, range "
                       f"[0, 100], duplicate removal, CompactPointField of the field",
           "cells": C, "out_cells": T, "out_points": U, "ms": round(ms, 4),
           "Gcells_per_s": round(C / (ms * 1e-3) / 1e9, 3), "alg_bytes": alg,
           "GBps": round(alg / (ms * 1e-3) / 1e9, 1)}
    res["frac_hbm_peak"] = round(res["GBps"] / HBM_PEAK_GBS, 4)
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        f_h = f.cpu().numpy()
        cb, counts = _omp_baseline(
            lambda th: orc.threshold_pipeline_omp(f_h, lo, hi, dedup=True, threads=th),
            C, "Gcells/s", "OpenMP restatement of Dax's OpenMP-adapter threshold sequence "
            "(count, scan, upper bounds, topology, point mask, stream compact, parallel "
            "sort+unique, lower bounds)", args.cpu_seconds)
        assert counts == (T, U)
        res["cpu_baseline"] = cb
    del f, out, vals
    torch.cuda.empty_cache()
    _progress(f"threshold {n}^3: {ms:.3f} ms, T={T}, U={U}")
    return res


def vtk_write_512(args):
    """The 512^3 benchmark contour (dedup) written as a VTK legacy POLYDATA
    file (the reference benchmark's PrintContentsToStream format) by the
    library's multi-threaded host writer, wall-clock from host arrays to a
    closed file under /tmp; beside it the oracle's ofstream restatement of
    the reference's writer on the same arrays (the files are compared)."""
    import tempfile

    import torch

    from daxtoolkit_amd import fields, io, mc
    n = 512
    dims = (n, n, n)
    field = fields.gaussian_mixture_torch(dims, fields.gaussian_mixture_params(0, dims))
    out = mc.forward(field, 0.5, normals=False)
    pts = out.vertices.cpu().numpy()
    conn = out.connectivity.reshape(-1, 3).cpu().numpy()
    del field, out
    torch.cuda.empty_cache()
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        a, b = f"{d}/lib.vtk", f"{d}/ref.vtk"
        t0 = time.perf_counter()
        io.write_vtk_polydata(a, pts, conn)
        sec = time.perf_counter() - t0
        size = os.path.getsize(a)
        res = {"workload": f"VTK legacy POLYDATA file of the {n}^3 benchmark contour "
                           f"({conn.shape[0]} triangles, {pts.shape[0]} points)",
               "bytes": size, "s": round(sec, 4), "MB_per_s": round(size / sec / 1e6, 1),
               "threads": os.cpu_count()}
        if not args.no_cpu_baseline:
            from oracle import oracle as orc
            t0 = time.perf_counter()
            orc.write_vtk(b, pts, conn, polydata=True)
            ref_sec = time.perf_counter() - t0
            with open(a, "rb") as fa, open(b, "rb") as fb:
                same = fa.read() == fb.read()
            res["identical_to_reference_stream"] = same
            res["cpu_baseline"] = {"value": round(size / ref_sec / 1e6, 1), "unit": "MB/s",
                                   "cores": 1, "kind": "port",
                                   "sample": f"the same arrays through an ofstream as the "
                                             f"reference writes them, {ref_sec:.2f} s"}
    _progress(f"vtk write: {sec:.3f} s, {size / 1e6:.1f} MB")
    return res


def hexahedra_256(args):
    """Marching cubes on the unstructured hexahedral grid of the 256^3
    uniform grid's voxels (MarchingCubesCount / MarchingCubesGenerate on
    UnstructuredGrid<Hexahedron>) on the 16-Gaussian field, iso 0.5, dedup
    (the general sort/unique/lower-bounds path); synchronous like the
    tetrahedra.  Beside it the oracle's serial restatement."""
    import torch

    from daxtoolkit_amd import fields, mc
    n = 256
    dims = (n, n, n)
    field = fields.gaussian_mixture_torch(dims, fields.gaussian_mixture_params(0, dims))
    hexes, pts = fields.voxel_cells_torch(dims, "hexahedron")
    f = field.reshape(-1)
    iso = 0.5
    out = mc.hexahedra(hexes, f, pts, iso)
    T, U = out.num_triangles, out.num_points
    reps = max(3, min(args.steps, 10))
    ms = _timed(lambda: mc.hexahedra(hexes, f, pts, iso), reps)
    C = int(hexes.shape[0])
    res = {"workload": f"marching cubes on an unstructured hexahedral grid (the {C} voxels of the "
                       f"{n}^3 grid as hexahedra), 16-Gaussian mixture (seed 0) + "
                       f"U(-0.01,0.01) noise, iso 0.5, dedup",
           "hexahedra": C, "triangles": T, "vertices": U, "ms": round(ms, 4),
           "Gcells_per_s": round(C / (ms * 1e-3) / 1e9, 3),
           "alg_bytes": _cells_alg_bytes(C, 8, T, U)}
    res["GBps"] = round(res["alg_bytes"] / (ms * 1e-3) / 1e9, 1)
    res["frac_hbm_peak"] = round(res["GBps"] / HBM_PEAK_GBS, 4)
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        hexes_h, pts_h, f_h = hexes.cpu().numpy(), pts.cpu().numpy(), f.cpu().numpy()
        cb, counts = _omp_baseline(
            lambda th: orc.unstructured_pipeline_omp(hexes_h, f_h, pts_h, iso, dedup=True,
                                                     threads=th),
            C, "Gcells/s", "OpenMP restatement of Dax's OpenMP-adapter marching-cubes sequence "
            "on hexahedra (classify, scan, upper/lower bounds, interpolation records, parallel "
            "sort+unique dedup, coordinates)", args.cpu_seconds)
        assert counts == (T, U)
        res["cpu_baseline"] = cb
    del hexes, pts, field
    torch.cuda.empty_cache()
    _progress(f"hexahedra {n}^3: {ms:.3f} ms, T={T}")
    return res


def large(args):
    import torch

    from daxtoolkit_amd import fields
    n = 1024
    t0 = time.perf_counter()
    field = fields.mandelbulb_torch(n)
    torch.cuda.synchronize()
    print(f"[bench] mandelbulb {n}^3 generated in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    a = argparse.Namespace(**vars(args))
    a.steps = max(3, min(args.steps, 10))
    r = single_volume(field, 0.5, "mandelbulb", a)
    out = {"workload": "1024^3 Mandelbulb escape-time field (power 8, <=10 iterations, bailout 2, "
                       "[-1.2,1.2]^3), iso 0.5, dedup, normals, edge list",
           "points": r["P"], "cells": r["C"], "triangles": r["T"], "vertices": r["U"],
           "fwd_ms": round(r["ms"], 4), "Gcells_per_s_fwd": round(r["C"] / (r["ms"] * 1e-3) / 1e9, 3),
           "tris_per_s": round(r["T"] / (r["ms"] * 1e-3), 1),
           "kernel_ms": [round(x, 4) for x in r["k_ms"]], "tuning": r["tuned"]}
    if "fb_ms" in r:
        out.update({"fwd_bwd_ms": round(r["fb_ms"], 4), "bwd_ms": round(r["bwd_ms"], 4),
                    "fwd_bwd_over_fwd": round(r["fb_ms"] / r["ms"], 3)})
    return out


def dispatcher_pipeline(n, reps=7):
    """The reference benchmark's timed region through the C++ dispatchers:
    the compiled program prints one JSON line with per-repetition seconds,
    each from the adapter's Timer<DeviceAdapterTagHip> around classify +
    generate into a fresh output grid, as the reference brackets them
    (Pipeline.h:110-132)."""
    import subprocess
    exe = ROOT / "tests" / "cpp" / "bin" / "BenchmarkPipelineMarchingCubesHip"
    if not exe.exists():
        return {"skipped": f"{exe.relative_to(ROOT)} is not built"}
    p = subprocess.run([str(exe), f"--size={n}", "--pipeline=2", f"--reps={reps}"],
                       capture_output=True, text=True, timeout=600)
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if p.returncode != 0 or not lines:
        return {"failed": p.returncode, "stderr": p.stderr[-500:]}
    d = json.loads(lines[-1])
    secs = sorted(d["seconds"][1:] or d["seconds"])  # the first repetition allocates
    med = statistics.median(secs)
    C = (n - 1) ** 3
    return {"workload": f"reference MC benchmark pipeline 2 (duplicate removal), {n}^3 grid, "
                        f"field |p| (DispatcherMapField<Magnitude>), iso 100",
            "triangles": d["triangles"], "points": d["points"],
            "median_ms": round(med * 1e3, 4), "min_ms": round(secs[0] * 1e3, 4),
            "Gcells_per_s": round(C / med / 1e9, 3), "reps": len(d["seconds"])}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported (the volume "
              "would be split into z-slabs with a 2-plane halo, one slab per GPU)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from daxtoolkit_amd import fields

    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(args.steps)
        return
    if args.only == "batch":
        print(json.dumps({"batched": batched(args)}), flush=True)
        return
    if args.only == "large":
        print(json.dumps({"large": large(args)}), flush=True)
        return
    if args.only == "diff":
        print(json.dumps({"fwd_bwd_256": differentiable_256(args)}), flush=True)
        return
    if args.only == "slice":
        print(json.dumps({"slice_512": slice_512(args)}), flush=True)
        return
    if args.only == "tets":
        print(json.dumps({"tetrahedra_128": tetrahedra_128(args)}), flush=True)
        return
    if args.only == "hex":
        print(json.dumps({"hexahedra_256": hexahedra_256(args)}), flush=True)
        return
    if args.only == "threshold":
        print(json.dumps({"threshold_512": threshold_512(args)}), flush=True)
        return
    if args.only == "vtk":
        print(json.dumps({"vtk_write_512": vtk_write_512(args)}), flush=True)
        return
    if args.only == "dispatch":
        print(json.dumps({"dispatcher_pipeline_512": dispatcher_pipeline(512)}), flush=True)
        return
    n = args.n
    dims = (n, n, n)
    iso = 0.5
    params = fields.gaussian_mixture_params(0, dims)
    field = fields.gaussian_mixture_torch(dims, params)
    torch.cuda.synchronize()
    tuning = [int(x) for x in args.tuning.split(",")] if args.tuning else None
    r = single_volume(field, iso, f"{n}^3", args, tuning)
    P, C, T, U, ms, k_ms, tuned = r["P"], r["C"], r["T"], r["U"], r["ms"], r["k_ms"], r["tuned"]

    W = ((n + 63) // 64) * n * n                       # row words
    b_fwd = 4 * P + 12 * U + 12 * U + 4 * 3 * T + 4 * U  # field + verts + normals + conn + edges
    fwd_gbs = b_fwd / (ms * 1e-3) / 1e9
    names = ["k_classify_v4", "k_count", "k_generate"]
    alg = [4 * P + P // 8,                 # field read once, one inside bit per point written
           P // 8 + 4 * W,                 # inside-bits read, one u32 word prefix per row word written
           12 * U + 12 * U + 4 * 3 * T + 4 * U + 8 * W]  # the outputs and two u32 bases per word
    traffic = _measured_traffic(n, tuned)
    kernels = {}
    for nm, t_k, a_k in zip(names, k_ms, alg):
        kernels[nm] = {"ms": round(t_k, 5), "alg_bytes": a_k, "GBps": round(a_k / (t_k * 1e-3) / 1e9, 1),
                       "frac": round(a_k / (t_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "traffic": traffic.get(nm) if traffic else None}
    dominant = names[max(range(3), key=lambda i: k_ms[i])]

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
        "wall_s": round(r["wall"], 4),
        "python_loop_ms_per_step": round(r["loop_ms"], 5),
        "python_loop_step_ms": {"min": round(min(r["steps_ms"]), 5),
                                "median": round(statistics.median(r["steps_ms"]), 5),
                                "max": round(max(r["steps_ms"]), 5)},
        "tuning": {"classify_unroll": tuned[0], "classify_vec": tuned[1], "scan_sub": tuned[2],
                   "gen_tile": tuned[3], "gen_xcd": tuned[4],
                   "autotune_ms": [round(x, 5) for x in r["tuned_ms"]]},
        "roofline": {"bound": "hbm", "scope": "forward graph: k_classify_v4, k_count (+ k_scan_tsums), k_generate",
                     "achieved": round(fwd_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fwd_gbs / HBM_PEAK_GBS, 4),
                     "traffic": (sum(traffic[k] for k in names)
                                 if traffic and all(traffic.get(k) for k in names) else None),
                     "traffic_source": TRAFFIC_NOTE,
                     "kernel_variants": list(_kernel_variants(n, tuned)),
                     "dominant_kernel": dominant, "kernels": kernels},
    }
    if "fb_ms" in r:
        line["fwd_bwd"] = {"ms": round(r["fb_ms"], 5), "bwd_ms": round(r["bwd_ms"], 5),
                           "fwd_bwd_over_fwd": round(r["fb_ms"] / ms, 3),
                           "Gcells_per_s": round(C / (r["fb_ms"] * 1e-3) / 1e9, 3)}
    if args.only != "fwd":
        if not args.no_bwd:
            line["fwd_bwd_256"] = differentiable_256(args)
        line["slice_512"] = slice_512(args)
        line["tetrahedra_128"] = tetrahedra_128(args)
        line["hexahedra_256"] = hexahedra_256(args)
        line["threshold_512"] = threshold_512(args)
        line["vtk_write_512"] = vtk_write_512(args)
        if args.batch > 0:
            line["batched"] = batched(args)
        if not args.no_large:
            line["large"] = large(args)
        line["dispatcher_pipeline_512"] = dispatcher_pipeline(512)
    if not args.no_cpu_baseline and args.only is None:
        f_host = field.cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(f_host, iso, C, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
