"""Benchmark of the 3photons event loop on one MI355X.

Workload (one step): e+e- -> 3 photons at the shipped valeurs, xoshiro256+
stream (seed_from_u64(12345)), 1e9 events = 1e5 batches of 10 000, batch
accumulators and their in-order fold left in HBM (tp_run_device).  The f64
path gives `value`; the mixed-precision path (f32 spinor/amplitude algebra)
is timed on 1e10 events of the same stream (1e6 batches) and reported beside
it with its per-quantity relative error against an f64 run of those events.  Also reported: the adjoint (2^24
points) and the learned proposal (training, then the same number of
weighted events, sigma and GPU-seconds against plain RAMBO at equal
error).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--events N] [--fe-events N]
                  [--no-mixed] [--no-adjoint] [--no-flow] [--flow-steps S]
                  [--flow-batch B] [--cpu-sample N] [--no-cpu]
  python bench.py --copy-calibration   # 3 plain 1 GiB device copies, for
                                       # scaling FETCH_SIZE / WRITE_SIZE
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
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
# Timed steps of each faster-evgen leg (1e8 events, about 12-15 ms each).
FE_STEPS = 3


# Host cores the measurement box allots to one GPU: the worker-pool size the
# baseline uses there, whatever the machine's own core count is.
HOST_CORE_SHARE = 16


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle() -> tuple[str, str]:
    """The oracle built for this host (make -C oracle native: -O3
    -march=native -fno-tree-slp-vectorize -ffp-contract=off), or the
    portable build if that fails."""
    import subprocess
    odir = os.path.join(ROOT, "oracle")
    try:
        # -B: always compile here, for this host's CPU
        subprocess.run(["make", "-B", "-s", "-C", odir, "native"], check=True, capture_output=True, timeout=300)
        return os.path.join(odir, "build", "native", "liboracle.so"), "gcc -O3 -march=native -fno-tree-slp-vectorize -ffp-contract=off"
    except (OSError, subprocess.SubprocessError) as e:
        return (os.path.join(odir, "build", "liboracle.so"),
                f"gcc -O3 -march=x86-64-v2 -ffp-contract=off (native build failed: {type(e).__name__})")


def cpu_baseline(sample_events: int, sample_1thread: int):
    """The oracle (C restatement of the reference CPU path) multi-threaded
    over batches like the reference's reproducible multi-threading mode
    (multi_threading.rs:16-75), on a bounded sample of the same workload,
    on the host's core share and on one thread.  Returns the JSON object and
    the oracle's accumulator of the sample (for the GPU's error against it)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle
    path, build = native_oracle()
    oracle = Oracle(path=path)
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(HOST_CORE_SHARE, affinity))
    cfg = oracle.config(VALEURS, sample_events)
    t0 = time.perf_counter()
    merged, _ = oracle.run(cfg, oracle.opts("xoshiro", threads=threads))
    dt = time.perf_counter() - t0
    cfg1 = oracle.config(VALEURS, sample_1thread)
    t0 = time.perf_counter()
    oracle.run(cfg1, oracle.opts("xoshiro", threads=1))
    dt1 = time.perf_counter() - t0
    out = {"value": sample_events / dt, "unit": "events/s", "cores": threads, "kind": "port",
           "sample": f"first {sample_events:.3g} events of the same xoshiro256+ f64 stream "
                     f"({(sample_events + 9999) // 10000} batches), {dt:.2f} s wall on {threads} threads",
           "value_1thread": sample_1thread / dt1,
           "sample_1thread": f"first {sample_1thread:.3g} events, {dt1:.2f} s wall on 1 thread",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "affinity_cpus": affinity,
           "core_share_note": f"the box allots {HOST_CORE_SHARE} host cores per GPU; the baseline uses "
                              f"min({HOST_CORE_SHARE}, affinity) threads",
           "build": build}
    return out, merged


def rel_errors(tp, cfg, acc, ref) -> dict:
    """Per-quantity relative error of run accumulator acc against ref (both
    finalised with the same host code): the res.data quantities, the |M|^2
    sums and second moments; the cancelling interference sums R_MX, I_MX
    against their statistical scale sqrt(N sum x^2)."""
    fa, fr = tp.finalize(cfg, acc), tp.finalize(cfg, ref)
    e = {q: abs(getattr(fa, q) - getattr(fr, q)) / abs(getattr(fr, q))
         for q in ("sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m")}
    n = cfg.num_events
    for k, name in enumerate(["A", "B+", "B-", "R_MX", "I_MX"]):
        if k < 3:
            e[f"spm2[{name}]"] = abs(acc.spm2[k] - ref.spm2[k]) / abs(ref.spm2[k])
        else:
            e[f"spm2[{name}]/stat"] = abs(acc.spm2[k] - ref.spm2[k]) / math.sqrt(n * ref.vars[k])
        e[f"vars[{name}]"] = abs(acc.vars[k] - ref.vars[k]) / abs(ref.vars[k])
    e["selected_events_diff"] = int(acc.selected_events) - int(ref.selected_events)
    return e


def traffic_from_profiles(name: str):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 --pmc
    summary profiles/<name>.md (its 'HBM bytes per launch:' line), if any."""
    path = os.path.join(ROOT, "profiles", f"{name}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("HBM bytes per launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def finite_or_none(x):
    """JSON has no NaN: a non-finite figure is printed as null."""
    return x if math.isfinite(x) else None


def json_safe(o):
    if isinstance(o, float):
        return finite_or_none(o)
    if isinstance(o, dict):
        return {k: json_safe(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [json_safe(v) for v in o]
    return o


def measure(tp, torch, eng, params, events, steps, warmup, per_batch, merged, kind=None, seeding=None):
    kind = tp.RNG_XOSHIRO if kind is None else kind
    seeding = tp.SEED_SKIPAHEAD if seeding is None else seeding

    def step():
        st = tp.rng_state(kind)
        eng.run_device(params, kind, seeding, st, events, per_batch.data_ptr(), merged.data_ptr())
        return eng.sync()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kernel_ms.append(step().kernel_ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    acc = tp.TpAcc.from_buffer_copy(bytes(merged.cpu().numpy().tobytes()))
    return steps * events / dt, dt, sum(kernel_ms) / len(kernel_ms), acc


def measure_adjoint(tp, torch, eng, params, n: int) -> dict:
    """Forward |M|^2*cut*w for the 5 contributions plus the hand-written
    gradients w.r.t. the 12 uniforms and 12 photon momentum components, for
    n points of the xoshiro256+ seed-42 stream, outputs left in HBM."""
    dev = torch.device("cuda:0")
    res = {"points": n, "stream": "xoshiro256+ seed_from_u64(42), 12 uniforms per point",
           "outputs_per_point": "5 values + 5x12 du + 5x12 dp"}
    F = torch.empty(n * 5 * 8, dtype=torch.uint8, device=dev)
    gu = torch.empty(n * 60 * 8, dtype=torch.uint8, device=dev)
    gp = torch.empty(n * 60 * 8, dtype=torch.uint8, device=dev)
    for prec, name in ((tp.PREC_F64, "f64"), (tp.PREC_F32, "f32")):
        st = tp.rng_state(tp.RNG_XOSHIRO, 42)
        eng.adjoint_device(params, 4096, prec, st, F.data_ptr(), gu.data_ptr(), gp.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.adjoint_device(params, n, prec, st, F.data_ptr(), gu.data_ptr(), gp.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k_ms = eng.sync().kernel_ms
        res[name] = {"points_per_s": n / dt, "wall_ms": dt * 1e3, "kernel_ms": k_ms}
    return res


def measure_flow(tp, eng, cfg, plain_value, plain_acc, events, steps, batch) -> dict:
    """Learned proposal: reverse-KL training on the GPU (steps x batch), then
    `events` weighted events from the 0.9 flow / 0.1 uniform mixture; sigma,
    its standard error and the GPU-seconds plain RAMBO needs for the same
    error, from the plain f64 run of the same size."""
    c = cfg.with_events(events)
    pc = tp.params_from_config(c)
    fl = tp.Flow(eng, init_seed=1)
    opts = tp.flow_default_opts(steps=steps, batch=batch)
    t0 = time.perf_counter()
    loss, st = fl.train(pc, opts)
    train_s = time.perf_counter() - t0
    nb = (events + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    fl.run(pc, opts.alpha, 99, min(events, 10 * tp.BATCH_SIZE))  # warm-up
    t0 = time.perf_counter()
    _, accf, sf = fl.run(pc, opts.alpha, 12345, events)
    run_s = time.perf_counter() - t0
    rf, rp = tp.finalize(c, accf), tp.finalize(c, plain_acc)
    se_f, se_p = rf.sigma * rf.prec, rp.sigma * rp.prec
    plain_s = events / plain_value
    # plain GPU-seconds to reach the flow's standard error (error ~ 1/sqrt(N))
    plain_equal = plain_s * (se_p / se_f) ** 2
    fl.close()
    return {
        "train": {"steps": steps, "batch": batch, "wall_s": train_s, "device_ms": st.device_ms,
                  "loss_first": float(loss[:max(1, steps // 20)].mean()),
                  "loss_last": float(loss[-max(1, steps // 20):].mean())},
        "events": events, "alpha": opts.alpha,
        "sigma_flow": rf.sigma, "se_flow": se_f, "sigma_plain": rp.sigma, "se_plain": se_p,
        "sigma_diff_in_se": (rf.sigma - rp.sigma) / (se_f ** 2 + se_p ** 2) ** 0.5,
        "variance_gain_per_event": (rp.prec / rf.prec) ** 2,
        "gpu_s_flow_run": run_s, "gpu_s_flow_kernels": sf.device_ms / 1e3, "gpu_s_plain_same_events": plain_s,
        "gpu_s_plain_at_flow_error": plain_equal,
        "gpu_s_ratio_excl_training": plain_equal / run_s,
        "gpu_s_ratio_incl_training": plain_equal / (run_s + train_s),
    }


def copy_calibration() -> int:
    """Three device-to-device copies of 1 GiB: the reference point that turns
    the FETCH_SIZE / WRITE_SIZE counters of a profiled run into bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"launches": 3, "bytes_read": n * 4, "bytes_written": n * 4}}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=1_000_000_000)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-events", type=int, default=10_000_000_000)
    ap.add_argument("--cpu-sample", type=int, default=100_000_000)
    ap.add_argument("--cpu-sample-1thread", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--fe-events", type=int, default=100_000_000)
    ap.add_argument("--adjoint-points", type=int, default=1 << 24)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--flow-steps", type=int, default=2000)
    ap.add_argument("--flow-batch", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); the event loop would "
              "shard by contiguous batch ranges, one range per GPU, but no multi-GPU run is part of it.",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration()

    import torch
    sys.path.insert(0, PKG)
    import trois_photons as tp

    cfg = tp.Configuration.load(VALEURS).with_events(args.events)
    nb = (args.events + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    nb_max = (max(args.events, 0 if args.no_mixed else args.mixed_events) + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    dev = torch.device("cuda:0")
    per_batch = torch.empty(nb_max * 104, dtype=torch.uint8, device=dev)
    merged = torch.empty(104, dtype=torch.uint8, device=dev)
    eng = tp.Engine(0)

    p64 = tp.params_from_config(cfg, precision=tp.PREC_F64)
    value, dt, k_ms, acc64 = measure(tp, torch, eng, p64, args.events, args.steps, args.warmup, per_batch, merged)
    achieved = F_ALG * args.events / (k_ms * 1e-3) / 1e12
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
        "dtype": "f64",
        "data": "synthetic (xoshiro256+ stream, shipped valeurs)",
        "config": {"workload": "e+e- -> 3 photons, RAMBO + cuts + helicity amplitudes, xoshiro256+ seed 12345, "
                               f"{args.events:.3g} events in {nb} batches of 10000, f64",
                   "events": args.events, "batches": nb, "rng": "xoshiro256+", "precision": "f64"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_from_profiles("fused_kernel_f64_xoshiro_1e9_traffic"),
                     "kernel": "fused_kernel<0,0> (f64, xoshiro)", "kernel_ms": k_ms, "flops_per_event": F_ALG},
    }
    if not args.no_mixed:
        # mixed precision on 1e10 events, against an f64 run of the same stream
        nm = args.mixed_events
        cm = cfg.with_events(nm)
        pm = tp.params_from_config(cm, precision=tp.PREC_MIXED)
        v32, dt32, k32, acc32 = measure(tp, torch, eng, pm, nm, args.steps, args.warmup, per_batch, merged)
        if nm == args.events:
            acc64m = acc64
        else:
            _, _, _, acc64m = measure(tp, torch, eng, tp.params_from_config(cm), nm, 1, 0, per_batch, merged)
        f64, f32 = tp.finalize(cm, acc64m), tp.finalize(cm, acc32)
        rel = {q: abs(getattr(f32, q) - getattr(f64, q)) / abs(getattr(f64, q))
               for q in ("sigma", "prec", "variance", "beta_min", "ss_p", "ss_m")}
        names = ["A", "B+", "B-", "R_MX", "I_MX"]
        for k in range(5):
            rel[f"spm2[{names[k]}]"] = abs(acc32.spm2[k] - acc64m.spm2[k]) / abs(acc64m.spm2[k])
            rel[f"vars[{names[k]}]"] = abs(acc32.vars[k] - acc64m.vars[k]) / abs(acc64m.vars[k])
        a32 = F_ALG * nm / (k32 * 1e-3) / 1e12
        out["fp32_mixed"] = {
            "value": v32, "unit": "events/s", "ms_per_step": dt32 * 1e3 / args.steps, "dtype": "f32+f64",
            "events": nm, "batches": (nm + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE,
            "selected_events": int(acc32.selected_events), "selected_events_f64": int(acc64m.selected_events),
            "rel_err_vs_f64": rel,
            "roofline": {"bound": "valu", "achieved": a32, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": a32 / FP32_PEAK_TFLOPS, "kernel": "fused_kernel<0,1> (mixed, xoshiro)",
                         "kernel_ms": k32, "flops_per_event": F_ALG},
        }
    if not args.no_variants:
        # the reference's default stream (RANF, expanded into HBM per chunk of
        # batches) and its all-f32 build (xoshiro128+, sums in event order)
        vr, dtr, kr, accr = measure(tp, torch, eng, p64, args.events, 2, 1, per_batch, merged, tp.RNG_RANF)
        p32 = tp.params_from_config(cfg, precision=tp.PREC_F32)
        v3, dt3, k3, acc3 = measure(tp, torch, eng, p32, args.events, 2, 1, per_batch, merged, tp.RNG_XOSHIRO128)
        f64r, f32r = tp.finalize(cfg, acc64), tp.finalize(cfg, acc3, f32=True)  # different streams
        # spinor.rs:40 takes sqrt(E + p_z): in f32 a photon along -z can round
        # to E < |p_z|, and that event's NaN reaches its batch's sums (about
        # 4.5 per 1e8 events, as in the reference's own f32 build)
        nb = (args.events + 9999) // 10000
        sums = per_batch[:nb * 104].view(torch.float64).reshape(nb, 13)[:, 1:]
        nan_batches = int((~torch.isfinite(sums).all(dim=1)).sum().item())
        # faster-evgen: event starts in the RANF stream depend on every earlier
        # draw (rejection re-rolls, round discards), so the stream is walked
        # first (on the device, segment walkers joined where they meet); timed
        # with that walk, on a bounded sample.  The reference's faster-evgen +
        # multi-threading build: 1e7 events in 0.87 s (VALIDATION.md:45-47)
        fe_events = min(args.events, args.fe_events)
        pfe = tp.params_from_config(cfg.with_events(fe_events), faster_evgen=True)
        vfe, dtfe, kfe, accfe = measure(tp, torch, eng, pfe, fe_events, FE_STEPS, 1, per_batch, merged, tp.RNG_RANF)
        # faster-evgen with faster-threading seeding, the reference's fastest
        # builds (VALIDATION.md:50-60: 1e7 events on a 4-core i7-4720HQ)
        fe_jump = {}
        for name, kind, prec, ref_s in (("ranf_f64", tp.RNG_RANF, tp.PREC_F64, 0.57),
                                        ("xoshiro_f64", tp.RNG_XOSHIRO, tp.PREC_F64, 0.59),
                                        ("ranf_f32", tp.RNG_RANF, tp.PREC_F32, 0.48),
                                        ("xoshiro128_f32", tp.RNG_XOSHIRO128, tp.PREC_F32, None)):
            pj = tp.params_from_config(cfg.with_events(fe_events), precision=prec, faster_evgen=True)
            vj, dtj, kj, accj = measure(tp, torch, eng, pj, fe_events, FE_STEPS, 1, per_batch, merged, kind, tp.SEED_JUMP)
            fe_jump[name] = {"value": vj, "unit": "events/s", "events": fe_events, "ms_per_step": dtj * 1e3 / FE_STEPS,
                             "kernels_ms": kj, "selected_events": int(accj.selected_events),
                             "reference_published_events_per_s": None if ref_s is None else 1e7 / ref_s}
        # faster-evgen with skip-ahead seeding on the other streams: xoshiro256+
        # (standard-random, no published figure), RANF in f32 (VALIDATION.md:41-43,
        # single-threaded) and xoshiro128+ in f32 (standard-random,f32), each
        # walked on the device as above
        fe_more = {}
        for name, kind, prec, ref_s in (("xoshiro_f64", tp.RNG_XOSHIRO, tp.PREC_F64, None),
                                        ("ranf_f32", tp.RNG_RANF, tp.PREC_F32, 2.11),
                                        ("xoshiro128_f32", tp.RNG_XOSHIRO128, tp.PREC_F32, None)):
            pj = tp.params_from_config(cfg.with_events(fe_events), precision=prec, faster_evgen=True)
            vj, dtj, kj, accj = measure(tp, torch, eng, pj, fe_events, FE_STEPS, 1, per_batch, merged, kind)
            fe_more[name] = {"value": vj, "unit": "events/s", "events": fe_events, "ms_per_step": dtj * 1e3 / FE_STEPS,
                             "kernels_ms": kj, "selected_events": int(accj.selected_events),
                             "reference_published_events_per_s": None if ref_s is None else 1e7 / ref_s}
        out["variants"] = {
            "ranf_f64": {"value": vr, "unit": "events/s", "ms_per_step": dtr * 1e3 / 2,
                         "kernels_ms": kr, "selected_events": int(accr.selected_events)},
            "all_f32_xoshiro128": {"value": v3, "unit": "events/s", "ms_per_step": dt3 * 1e3 / 2, "kernels_ms": k3,
                                   "selected_events": int(acc3.selected_events),
                                   "sigma_pb": finite_or_none(f32r.sigma), "prec": finite_or_none(f32r.prec),
                                   "nonfinite_batches": nan_batches,
                                   "sigma_pb_f64_xoshiro256": f64r.sigma},
            "faster_evgen_f64": {"value": vfe, "unit": "events/s", "events": fe_events, "ms_per_step": dtfe * 1e3 / FE_STEPS,
                                 "kernels_ms": kfe, "selected_events": int(accfe.selected_events),
                                 "includes": "walk of the RANF stream for the event starts (device)",
                                 "reference_published_events_per_s": 1e7 / 0.87},
            "faster_evgen_skipahead": fe_more,
            "faster_evgen_faster_threading": fe_jump,
        }
    if not args.no_adjoint:
        out["adjoint"] = measure_adjoint(tp, torch, eng, p64, args.adjoint_points)
    if not args.no_flow:
        out["flow"] = measure_flow(tp, eng, cfg, value, acc64, args.events, args.flow_steps, args.flow_batch)
    if not args.no_cpu:
        out["cpu_baseline"], oacc = cpu_baseline(args.cpu_sample, args.cpu_sample_1thread)
        # the FP64 oracle's result on the sample against the GPU's f64 and
        # mixed runs of the same events
        ns = args.cpu_sample
        cs = cfg.with_events(ns)
        ref = tp.TpAcc.from_buffer_copy(bytes(oacc))
        accs = {}
        for name, prec in (("f64", tp.PREC_F64), ("mixed", tp.PREC_MIXED)):
            _, _, _, accs[name] = measure(tp, torch, eng, tp.params_from_config(cs, precision=prec), ns, 1, 0,
                                          per_batch, merged)
        out["rel_err_vs_oracle"] = {"events": ns, "f64": rel_errors(tp, cs, accs["f64"], ref)}
        if "fp32_mixed" in out:
            out["fp32_mixed"]["rel_err_vs_oracle"] = rel_errors(tp, cs, accs["mixed"], ref)
            out["fp32_mixed"]["rel_err_vs_oracle_events"] = ns
    print(json.dumps(json_safe(out), allow_nan=False))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
