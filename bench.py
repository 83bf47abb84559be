#!/usr/bin/env python3
"""Benchmark of the sparse solver hot path on one MI355X.

Headline (one JSON line on stdout): FP64 CSR SpMV throughput in GB/s of
algorithmic traffic on the 3-D 7-point Laplacian 128^3 (n = 2,097,152,
nnz = 14,581,760), inputs resident in HBM.  A "step" is one SpMV y = A x.

Algorithmic bytes per SpMV (value size V, int32 indices):
    B = nnz (V + 4) + 4 (n + 1) + 2 n V      (A once, x once, y once)

The line also carries
  * roofline: the SpMV kernel's average duration from HIP events on the
    stream it runs on, against MI355X's 8 TB/s HBM peak;
  * stream_copy_GBps: a device-to-device copy measured in the same run;
  * cpu_baseline: the oracle's OpenMP CSR SpMV (restated from
    omp/matrix/csr_kernels.cpp) on the host cores, same matrix;
  * secondary: further measurements of the path (FP32, SELL-P, power-law).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def spmv_bytes(n, nnz, vbytes):
    return nnz * (vbytes + 4) + 4 * (n + 1) + 2 * n * vbytes


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_kernel_loop(fn, steps, warmup, stream):
    """Runs fn() `warmup` times untimed, then `steps` times between HIP
    events recorded on `stream`; returns (wall_ms_per_step, event_ms_per_step)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, start.elapsed_time(stop) / steps


def stream_copy_gbps(dev):
    import torch
    n = 1 << 27  # 1 GiB of doubles
    a = torch.empty(n, dtype=torch.float64, device=dev)
    a.fill_(1.0)
    b = torch.empty_like(a)
    s = torch.cuda.current_stream(dev)
    _, ms = time_kernel_loop(lambda: b.copy_(a), 20, 5, s)
    del a, b
    return 2 * n * 8 / (ms * 1e-3) / 1e9


def cpu_spmv_baseline(rp, ci, v, x, budget_s=12.0):
    """Oracle OpenMP SpMV on the host (the CPU baseline, not a target)."""
    import numpy as np
    import oracle as ora
    ora.csr_spmv(rp, ci, v, x, omp=True)
    reps = 0
    t0 = time.perf_counter()
    while True:
        ora.csr_spmv(rp, ci, v, x, omp=True)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= 200:
            break
    per = el / reps
    n = len(rp) - 1
    return {
        "value": round(spmv_bytes(n, int(rp[-1]), v.itemsize) / per / 1e9, 3),
        "unit": "GB/s",
        "cores": ora.omp_threads(),
        "kind": "port",
        "sample": f"{reps} OpenMP CSR SpMVs of the same 128^3 Laplacian "
                  f"({per * 1e3:.2f} ms each, {el:.1f} s total)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="skip secondary measurements")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1 "
                                   "(multi-GPU sharding is a design note, not built)"}))
        return 2

    import numpy as np
    import torch
    from ginkgo_amd import device as D
    from ginkgo_amd import generators as gen

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)

    log("[bench] generating 128^3 7-pt Laplacian")
    rp, ci, v = gen.laplace3d(128)
    n, nnz = len(rp) - 1, int(rp[-1])
    rng = np.random.default_rng(2)
    x_h = rng.uniform(-1, 1, n)
    A = D.DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    x = torch.as_tensor(x_h).to(dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    log(f"[bench] n={n} nnz={nnz} strategy={A.strategy}")

    # correctness guard on the measured kernel
    A.spmv(x, y)
    torch.cuda.synchronize()

    log("[bench] timing FP64 CSR SpMV")
    wall_ms, ev_ms = time_kernel_loop(lambda: A.spmv(x, y), args.steps, args.warmup, stream)
    b64 = spmv_bytes(n, nnz, 8)
    value = b64 / (wall_ms * 1e-3) / 1e9
    achieved = b64 / (ev_ms * 1e-3) / 1e9

    stream_bw = stream_copy_gbps(dev)
    result = {
        "metric": "SpMV HBM GB/s (FP64 CSR, 3D 7-pt Laplacian 128^3)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, x ~ U[-1,1] seed 2)",
        "config": {"workload": "csr_spmv laplace3d_7pt 128^3", "n": n, "nnz": nnz,
                   "index": "int32", "strategy": A.strategy, "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": None, "kernel_avg_us": round(ev_ms * 1e3, 3),
                     "algorithmic_bytes": b64},
        "stream_copy_GBps": round(stream_bw, 1),
        "frac_of_stream_copy": round(achieved / stream_bw, 4),
    }

    secondary = {}
    if not args.quick:
        log("[bench] secondary: FP32 CSR, SELL-P, power-law")
        A32 = D.DeviceCsr(rp, ci, v.astype(np.float32), strategy="classical", device=dev)
        x32 = x.float()
        y32 = torch.empty(n, dtype=torch.float32, device=dev)
        _, ms = time_kernel_loop(lambda: A32.spmv(x32, y32), args.steps, args.warmup, stream)
        secondary["csr_f32_laplace128_GBps"] = round(spmv_bytes(n, nnz, 4) / (ms * 1e-3) / 1e9, 1)
        S = D.DeviceSellp(A, 64, 1)
        _, ms = time_kernel_loop(lambda: S.spmv(x, y), args.steps, args.warmup, stream)
        secondary["sellp_f64_laplace128_GBps"] = round(b64 / (ms * 1e-3) / 1e9, 1)
        del S, A32
        prp, pci, pv = gen.powerlaw(1_000_000, seed=3)
        pn, pnnz = len(prp) - 1, int(prp[-1])
        P = D.DeviceCsr(prp, pci, pv, strategy="load_balance", device=dev)
        px = torch.as_tensor(np.random.default_rng(3).uniform(-1, 1, pn)).to(dev)
        py = torch.empty(pn, dtype=torch.float64, device=dev)
        _, ms = time_kernel_loop(lambda: P.spmv(px, py), args.steps, args.warmup, stream)
        secondary["csr_f64_powerlaw1M_GBps"] = round(spmv_bytes(pn, pnnz, 8) / (ms * 1e-3) / 1e9, 1)
        secondary["powerlaw_nnz"] = pnnz
        del P
    if secondary:
        result["secondary"] = secondary

    if not args.no_cpu:
        log("[bench] CPU baseline (oracle OpenMP SpMV)")
        result["cpu_baseline"] = cpu_spmv_baseline(rp, ci, v, x_h.reshape(-1, 1))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
