#!/usr/bin/env python3
"""Benchmark of the ray-traced AO path on one MI355X.

Workload (a step is one frame): the seeded ~1M-triangle scene (2,000 boxes and
icospheres over a 40 x 40 ground quad), G-buffer at 1920 x 1080 from the host rasterizer
(uploaded once, resident in HBM before timing), then per frame on the GPU: PLOC BVH
build, ray-traced AO at 64 rays/pixel with radius 1.0, SSAO at 24 samples, the
reconcile's error map, and the reconcile's float64 sum on the host.

value = AO rays per second of whole frames (covered pixels x spp x steps / timed wall
time). Also reported: the trace kernel's own ray rate, per-kernel milliseconds, the trace
kernel's roofline from a counting run (node and triangle fetches), and the CPU baseline:
the reference's own computeRTAO (oracle/_ref, built from its sources) on a row sample of
the same frame, timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calib-copy MiB]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "raw-native_amd"))

METRIC = "AO rays/s + ms/frame at 1080p×64spp; learned-AO ms/frame at RMSE≤0.02 vs RT"
HBM_PEAK_GBS = 8000.0


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def cpu_baseline(tris, pos, nrm, mask, spp, radius, row_step):
    """The reference's computeRTAO (its own sources, oracle/_ref) on every row_step-th row."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    sub = slice(0, pos.shape[0], row_step)
    p2, n2, m2 = [np.ascontiguousarray(x[sub]) for x in (pos, nrm, mask)]
    ao = np.zeros(m2.shape, np.float32)
    if O.ref_available():
        R = O.ref()
        kind = "reference"
        t0 = time.perf_counter()
        acc = R.ref_accel_build(tris, tris.shape[0])
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        R.ref_rtao_accel(acc, p2, n2, m2, m2.shape[1], m2.shape[0], spp, radius, threads, ao)
        trace_s = time.perf_counter() - t0
        R.ref_accel_free(acc)
    else:
        kind = "port"
        t0 = time.perf_counter()
        b = O.OracleBvh(tris)
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ao = b.rtao(p2, n2, m2, spp, radius, threads=threads)
        trace_s = time.perf_counter() - t0
    rays = int(m2.sum()) * spp
    return {"value": rays / trace_s, "unit": "rays/s", "cores": threads, "kind": kind,
            "sample": "computeRTAO on every %dth row of the same frame (%d rays, %.2f s); BVH build (1 thread) %.2f s"
                      % (row_step, rays, trace_s, build_s),
            "build_s": build_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--radius", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-row-step", type=int, default=4)
    ap.add_argument("--calib-copy", type=int, default=0, help="also run a plain device copy of this many MiB "
                    "(the known-size kernel that scales the HBM counters)")
    ap.add_argument("--kernel", choices=["wide", "binary"], default="wide", help="traversal form")
    ap.add_argument("--tri-q", type=int, default=0, help="wide kernel's triangle-phase threshold (1..64; 0: default)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the AO path's multi-GPU form (replicated BVH, row-split frame) is "
              "not built in this tree; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import raw_ao

    W, H, spp, radius = a.width, a.height, a.spp, a.radius
    spec = raw_ao.host.SCENE_1M
    log("scene generation")
    vpos, vnrm, idx, nspheres = raw_ao.host.scene_seeded(spec)
    tris = raw_ao.host.mesh_tris(vpos, idx)
    eye = raw_ao.host.orbit_camera(1)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9,
                                              float(np.float32(W) / np.float32(H)), 0.1, 200.0, W, H)
    covered = int(mask.sum())
    log("%d triangles, %dx%d, %d covered pixels" % (tris.shape[0], W, H, covered))

    dev = torch.device("cuda:0")
    d_tris = torch.from_numpy(tris).to(dev)
    d_pos = torch.from_numpy(pos).to(dev)
    d_nrm = torch.from_numpy(nrm).to(dev)
    d_mask = torch.from_numpy(mask).to(dev)
    d_ao = torch.empty((H, W), dtype=torch.float32, device=dev)
    d_ss = torch.empty_like(d_ao)
    d_err = torch.empty_like(d_ao)
    d_max = torch.zeros(4, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ctx = raw_ao.Context(0)
    ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_BINARY if a.kernel == "binary" else raw_ao.TRACE_WIDE)
    if a.tri_q:
        ctx.set_option(raw_ao.OPT_TRI_PHASE, a.tri_q)

    host_ms = []

    def frame():
        ctx.build_ploc_dev(d_tris, 16, stream)
        ctx.compute_rtao_dev(d_pos, d_nrm, d_mask, W, H, spp, radius, d_ao, stream)
        ctx.compute_ssao_dev(d_pos, d_nrm, d_mask, W, H, 24, radius, d_ss, stream)
        ctx.error_map_dev(d_ss, d_ao, d_mask, W, H, d_err, d_max, stream)
        err = d_err.cpu()   # synchronizes the stream
        t0 = time.perf_counter()
        rmse, px = raw_ao.host.rmse_of(err.numpy(), mask)
        host_ms.append((time.perf_counter() - t0) * 1e3)
        return rmse, px

    for i in range(a.warmup):
        frame()
    log("warmup done")
    # counting launch (outside the timed region): node visits and triangle tests per frame
    ctx.count_next_trace(True)
    ctx.compute_rtao_dev(d_pos, d_nrm, d_mask, W, H, spp, radius, d_ao, stream)
    torch.cuda.synchronize()
    ctx.count_next_trace(False)
    # read the counters through a host-pointer call's bookkeeping
    visits, tests, rays_counted = _counts(ctx, raw_ao)
    log("counting run: %d rays, %.2f visits/ray, %.2f tests/ray" % (rays_counted, visits / max(1, rays_counted),
                                                                     tests / max(1, rays_counted)))
    ms = {"build": [], "trace": [], "ssao": [], "error": []}
    host_ms.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        rmse, px = frame()
        m = ctx.last_ms()
        for k in ms:
            ms[k].append(m[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log("timed %d frames in %.3f s" % (a.steps, wall))

    rays = covered * spp
    value = rays * a.steps / wall
    trace_ms = float(np.mean(ms["trace"]))
    node_bytes, tri_bytes, rec = ctx.bvh_bytes()
    alg_bytes = visits * rec + tests * 48 + covered * 24 + W * H * (1 + 4)
    achieved = alg_bytes / (trace_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "rays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "seeded %d-triangle scene (%d icospheres), PLOC build + ray-traced AO %dx%d at %d spp, "
                               "radius %.1f + SSAO 24 + reconcile, one frame per step" % (tris.shape[0], nspheres, W, H, spp,
                                                                                          radius),
                   "triangles": int(tris.shape[0]), "width": W, "height": H, "spp": spp, "radius": radius,
                   "covered_pixels": covered, "rays_per_frame": rays, "parallelism": "single GPU"},
        "trace_rays_per_s": rays / (trace_ms * 1e-3),
        "ms_frame_parts": {"ploc_build": float(np.mean(ms["build"])), "ao_trace": trace_ms,
                           "ssao": float(np.mean(ms["ssao"])), "error_map": float(np.mean(ms["error"])),
                           "host_reconcile_sum": float(np.mean(host_ms))},
        "rmse_ssao_vs_rt": float(rmse), "reconcile_pixels": px,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_trace_ao", "algorithmic_bytes_per_launch": alg_bytes,
                     "node_visits": visits, "triangle_tests": tests},
    }
    if a.calib_copy:
        n = a.calib_copy * (1 << 20) // 4
        x = torch.empty(n, dtype=torch.float32, device=dev).fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        out["calib_copy_mib"] = a.calib_copy
    if not a.no_cpu:
        log("CPU baseline")
        cb = cpu_baseline(tris, pos, nrm, mask, spp, radius, a.cpu_row_step)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["vs_cpu_trace"] = out["trace_rays_per_s"] / cb["value"]
    ctx.close()
    print(json.dumps(out))


def _counts(ctx, raw_ao):
    v, t, r = ctx.trace_counts()
    return v, t, r


if __name__ == "__main__":
    main()
