"""Benchmark of the 3photons event loop on one MI355X.

Workload (one step): e+e- -> 3 photons at the shipped valeurs, xoshiro256+
stream (seed_from_u64(12345)), f64, 1e9 events = 1e5 batches of 10 000, with
batch accumulators and their in-order fold left in HBM (tp_run_device).
Prints ONE JSON line (metric/value/... plus roofline and cpu_baseline).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--events N]
                  [--precision f64|mixed] [--cpu-sample N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "3photons-rust_amd")
VALEURS = os.path.join(ROOT, "tests", "golden", "valeurs")

METRIC = "events/s on 1 MI355X (FP64, FP32); GPU-s to equal σ error: flow vs RAMBO"
# Algorithmic work per generated event of the reference as written (each +,
# -, *, /, sqrt, sin, cos, ln = 1), frozen from oracle/opcount over 2e6
# events of this stream: 276.1 add/sub + 539.4 mul + 19.6 div + 8.2 sqrt +
# 9 transcendental (pass rate 0.708).
F_ALG = 852.3
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
HBM_PEAK_GBS = 8000.0


def cpu_baseline(sample_events: int) -> dict:
    """The oracle (C restatement of the reference CPU path) on the host's
    cores, multi-threaded over batches like the reference's reproducible
    multi-threading mode, on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle
    oracle = Oracle()
    threads = min(16, os.cpu_count() or 1)
    cfg = oracle.config(VALEURS, sample_events)
    opts = oracle.opts("xoshiro", threads=threads)
    t0 = time.perf_counter()
    oracle.run(cfg, opts)
    dt = time.perf_counter() - t0
    return {"value": sample_events / dt, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_events:.3g} events of the same xoshiro256+ f64 stream "
                      f"({(sample_events + 9999) // 10000} batches), oracle/liboracle.so, {dt:.2f} s wall"}


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("bytes_per_launch")
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=1_000_000_000)
    ap.add_argument("--precision", choices=["f64", "mixed"], default="f64")
    ap.add_argument("--cpu-sample", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); the event loop shards by "
              "batch range across GPUs as described in DESIGN.md, but no multi-GPU run is part of it.",
              file=sys.stderr)
        return 2

    import torch
    sys.path.insert(0, PKG)
    import trois_photons as tp

    cfg = tp.Configuration.load(VALEURS).with_events(args.events)
    prec = tp.PREC_F64 if args.precision == "f64" else tp.PREC_MIXED
    params = tp.params_from_config(cfg, precision=prec)
    kind = tp.RNG_XOSHIRO
    nb = (args.events + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    dev = torch.device("cuda:0")
    per_batch = torch.empty(nb * 104, dtype=torch.uint8, device=dev)
    merged = torch.empty(104, dtype=torch.uint8, device=dev)
    eng = tp.Engine(0)

    def step():
        st = tp.rng_state(kind)
        eng.run_device(params, kind, tp.SEED_SKIPAHEAD, st, args.events, per_batch.data_ptr(), merged.data_ptr())
        return eng.sync()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = step()
        kernel_ms.append(s.kernel_ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    value = args.steps * args.events / dt
    k_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = F_ALG * args.events / (k_ms * 1e-3) / 1e12
    peak = FP64_PEAK_TFLOPS if prec == tp.PREC_F64 else FP32_PEAK_TFLOPS
    kname = "fused_f64_xoshiro" if prec == tp.PREC_F64 else "fused_mixed_xoshiro"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == tp.PREC_F64 else "f32+f64",
        "data": "synthetic (xoshiro256+ stream, shipped valeurs)",
        "config": {"workload": "e+e- -> 3 photons, RAMBO + cuts + helicity amplitudes, xoshiro256+ seed 12345, "
                               f"{args.events:.3g} events in {nb} batches of 10000, {args.precision}",
                   "events": args.events, "batches": nb, "rng": "xoshiro256+", "precision": args.precision},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic_from_profiles(kname),
                     "kernel": kname, "kernel_ms": k_ms, "flops_per_event": F_ALG},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
