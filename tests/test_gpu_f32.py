"""The all-f32 mode on the GPU (the reference's `f32` feature): RANF values
as (n as f32) * 1e-9 and the xoshiro128+ stream, the per-event path in f32
with the reference's operations, batch sums in event order, f32 fold and
finalize, against the f32 oracle batch by batch and against the reference's
shipped f32 outputs at N = 1e7."""
import json
import math

import numpy as np
import pytest

import trois_photons as tp
from conftest import FE_EDGE_BATCHES, FE_TWO_CHUNKS, golden_path
from numdiff import printed_match
from resdata import golden_numbers, printed_values

pytestmark = pytest.mark.gpu

KIND = {"ranf": tp.RNG_RANF, "xoshiro": tp.RNG_XOSHIRO128}


@pytest.fixture(scope="module")
def engine():
    print("creating engine on cuda:0", flush=True)
    with tp.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def cfg(valeurs):
    return tp.Configuration.load(valeurs)


def close_f32(a, b, n_events, limit=2e-5):
    """Selected counts exact; sums to the f32 rounding that one-ulp
    differences of sinf/cosf/logf (ocml against glibc) leave (limit: see
    F32_CONDITIONING_LIMIT for the batches that hold an ill-conditioned
    event)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert a[0] == b[0], (a[0], b[0], n_events)
    assert np.array_equal(np.isnan(a), np.isnan(b)), (a, b, n_events)  # NaN batches: see test_f32_spinor_nan_*
    with np.errstate(invalid="ignore"):
        rel = np.abs(a[1:] - b[1:]) / np.maximum(np.abs(b[1:]), 1e-300)
        # I_MX (index 5 of the sums) cancels: compare it to the size of its terms
        rel[4] = abs(a[5] - b[5]) / max(abs(b[10]) ** 0.5 * max(b[0], 1) ** 0.5, 1e-300)
    rel = np.where(np.isnan(b[1:]), 0.0, rel)
    worst = float(rel.max()) if len(rel) else 0.0
    assert rel.max() <= limit, (n_events, rel)
    return worst


def conditioning_state():
    """The RANF state before the f32 faster-evgen event that f32 arithmetic
    gets percents away from f64 (tests/golden/f32_conditioning_event.json,
    pinned against both oracles by test_oracle_f32_conditioning_event), and
    the batch that holds it."""
    with open(golden_path("f32_conditioning_event.json")) as f:
        d = json.load(f)
    st = tp.rng_state(tp.RNG_RANF)
    for i, x in enumerate(d["ranf_numbers"]):
        st.numbers[i] = x
    st.index = d["ranf_index"]
    return st, d["batch"]


# A batch holding that event differs from the f32 oracle by up to one event's
# share of its sums (the event's own sums differ by percents): 9.3e-5 measured.
# Such events come about once in 1e8 on each stream.
F32_CONDITIONING_LIMIT = 5e-4


def close_f32_batches(per, oper):
    """Every batch of a run of whole batches against the f32 oracle's: counts
    exact in all; sums within 2e-5 in all but one batch in 4000, which stay
    within F32_CONDITIONING_LIMIT.  Returns those batches."""
    loose = []
    for b in range(len(per)):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        if close_f32(a, oper[b].as_array(), 10000, F32_CONDITIONING_LIMIT) > 2e-5:
            loose.append(b)
    print(f"  batches beyond 2e-5 (ill-conditioned events): {loose}")
    assert len(loose) <= max(1, len(per) // 4000), loose
    return loose


@pytest.mark.parametrize("rng", ["ranf", "xoshiro"])
@pytest.mark.parametrize("jump", [False, True])
@pytest.mark.parametrize("n_events", [1, 9999, 35017])
def test_f32_per_batch_matches_oracle(engine, cfg, oracle32, valeurs, rng, jump, n_events):
    print(f"f32 run rng={rng} jump={jump} N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32)
    st = tp.rng_state(KIND[rng])
    per, merged, _ = engine.run(params, KIND[rng], tp.SEED_JUMP if jump else tp.SEED_SKIPAHEAD, st, n_events)
    ocfg = oracle32.config(valeurs, n_events)
    om, oper = oracle32.run(ocfg, oracle32.opts(rng, faster_threading=jump, threads=4), per_batch=True)
    worst = 0.0
    for b in range((n_events + 9999) // 10000):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        worst = max(worst, close_f32(a, oper[b].as_array(), n_events))
        assert all(float(np.float32(x)) == x for x in a[1:]), "f32 values"
    close_f32(merged.as_array(), om.as_array(), n_events)
    print(f"  worst relative difference of a batch sum: {worst:.2e}")


def test_f32_simulate_batch_advances_the_stream(engine, cfg, oracle32, valeurs):
    """Closure calls of every size up to the 640 000-event maximum: a lane's
    xoshiro128+ draw offset inside one call reaches 12 * 640 000 > 2^22."""
    params = tp.params_from_config(cfg, precision=tp.PREC_F32)
    st = tp.rng_state(tp.RNG_XOSHIRO128)
    g = oracle32.rng("xoshiro")
    ocfg = oracle32.config(valeurs)
    opts = oracle32.opts("xoshiro")
    for n in (10000, 7, 1, 23456, 100000, 640000):
        print(f"f32 simulate_batch n={n}", flush=True)
        acc = engine.simulate_batch(params, tp.RNG_XOSHIRO128, st, n)
        close_f32(acc.as_array(), oracle32.simulate_batch(ocfg, opts, g, n).as_array(), n)
        assert list(st.s) == list(g.xs32)


@pytest.mark.parametrize("name,rng", [("res.data-features_f32", "ranf"),
                                      ("res.data-features_standard-random,f32", "xoshiro")])
def test_f32_full_run_reproduces_golden(engine, cfg, name, rng):
    """N = 1e7: every printed field of the reference's f32 res.data (CI
    compares these files with numdiff -a 1.1e-8, ci.yml:180-183)."""
    print(f"golden {name}", flush=True)
    res = tp.run_simulation(cfg, rng=rng, precision=tp.PREC_F32, engine=engine)
    golden = golden_numbers(open(golden_path(name)).read())
    vals = printed_values(cfg, res.final, f32=True)
    assert len(vals) == len(golden)
    bad = [(i, v, g) for i, (v, g) in enumerate(zip(vals, golden))
           if not (math.isnan(v) and g == "NaN") and not (abs(v - float(g)) <= 1.1e-8 or printed_match(v, g))]
    assert not bad, bad


def test_f32_rejects_mismatched_streams(engine, cfg):
    p32 = tp.params_from_config(cfg.with_events(10), precision=tp.PREC_F32)
    with pytest.raises(tp.TpError, match="xoshiro128"):
        engine.run(p32, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), 10)
    p64 = tp.params_from_config(cfg.with_events(10))
    with pytest.raises(tp.TpError, match="xoshiro128"):
        engine.run(p64, tp.RNG_XOSHIRO128, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO128), 10)


def test_f32_cli_end_to_end(tmp_path, valeurs):
    """trois_photons --f32: the reference's f32 res.data and stdout."""
    import os
    import shutil
    import subprocess
    from conftest import PKG
    from numdiff import compare_texts
    exe = os.path.join(PKG, "build", "trois_photons")
    shutil.copy(valeurs, tmp_path / "valeurs")
    print("cli --f32 run", flush=True)
    r = subprocess.run([exe, "--f32"], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    golden = open(golden_path("res.data-features_f32")).read()
    assert compare_texts(golden, (tmp_path / "res.data").read_text(), digits_slack=1.5) == []
    # the report also prints alpha0 = I_MX / 2 to 6 digits: the I_MX sum cancels (printed relative
    # error ~3), so its f32 value carries the f32 sums' last ulps amplified to ~1e-5
    assert compare_texts(open(golden_path("stdout.log-features_f32")).read(), r.stdout, rel=5e-5,
                         digits_slack=1.5) == []


@pytest.mark.parametrize("batch", [1158, 1880])
def test_f32_spinor_nan_batches_match_oracle(engine, cfg, oracle32, valeurs, batch):
    """Batches of the f32 xoshiro128+ stream that hold one event with
    E + p_z < 0 after f32 rounding (spinor.rs:40 takes its square root): the
    GPU gives NaN sums where the oracle does, the same selected count, and the
    neighbouring batches stay finite and match."""
    print(f"f32 NaN batch {batch} and neighbours", flush=True)
    params = tp.params_from_config(cfg, precision=tp.PREC_F32)
    st = tp.rng_state(tp.RNG_XOSHIRO128)
    tp.rng_skip(tp.RNG_XOSHIRO128, st, 120000 * (batch - 1))
    g = oracle32.rng("xoshiro")
    oracle32.skip(g, 120000 * (batch - 1))
    ocfg, opts = oracle32.config(valeurs), oracle32.opts("xoshiro")
    for b in (batch - 1, batch, batch + 1):
        a = engine.simulate_batch(params, tp.RNG_XOSHIRO128, st, 10000).as_array()
        o = oracle32.simulate_batch(ocfg, opts, g, 10000).as_array()
        assert list(st.s) == list(g.xs32)
        if b == batch:
            assert a[0] == o[0] and np.isnan(a[1:]).all() and np.isnan(o[1:]).all(), (a, o)
        else:
            assert np.isfinite(a).all()
            close_f32(a, o, 10000)


def test_f32_grid_stride_batches_match_oracle(engine, cfg, oracle32, valeurs):
    """N = 1e8 (10 000 batches, more than the launch's workgroups): each lane
    reaches its later batches by the grid-stride matrix of xoshiro128+ instead
    of a skip-ahead from the run's start.  The last batches, and one early
    one, against the f32 oracle started at those batches."""
    n_events = 100_000_000
    print(f"f32 run N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32)
    per, _, _ = engine.run(params, tp.RNG_XOSHIRO128, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO128), n_events)
    ocfg, opts = oracle32.config(valeurs, n_events), oracle32.opts("xoshiro")  # the weight carries 1/N
    nb = n_events // 10000
    for b in (7, nb - 3, nb - 2, nb - 1):
        g = oracle32.rng("xoshiro")
        oracle32.skip(g, 120000 * b)
        o = oracle32.simulate_batch(ocfg, opts, g, 10000).as_array()
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        close_f32(a, o, 10000)


def test_f32_ranf_batches_past_the_first_chunk_match_oracle(engine, cfg, oracle32, valeurs):
    """RANF in the all-f32 mode at N = 7e7 (7000 batches): the stream is
    expanded and the events and sums run per chunk of 6144 batches; the
    batches at the chunk edge and the last one against the f32 oracle
    started at those batches."""
    n_events = 70_000_000
    print(f"f32 RANF run N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32)
    per, _, _ = engine.run(params, tp.RNG_RANF, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_RANF), n_events)
    ocfg, opts = oracle32.config(valeurs, n_events), oracle32.opts("ranf")  # the weight carries 1/N
    for b in (0, 6143, 6144, 6999):
        g = oracle32.rng("ranf")
        oracle32.skip(g, 120000 * b)
        o = oracle32.simulate_batch(ocfg, opts, g, 10000).as_array()
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        if np.isfinite(o).all():
            close_f32(a, o, 10000)
        else:  # a batch whose f32 spinor products overflow: the same NaN sums on both sides
            assert a[0] == o[0] and np.array_equal(np.isnan(a), np.isnan(o)), (b, a, o)


@pytest.mark.parametrize("prec", [tp.PREC_F32, tp.PREC_F64])
def test_device_fold_equals_host_fold(engine, cfg, prec):
    """merge_kernel<float> (all-f32 runs) and merge_kernel<double>: the
    device's in-order fold of 2 345 batches (ten LDS tiles, the last partial)
    is the host's left fold, bit for bit."""
    import torch
    n = 23_450_000
    kind = tp.RNG_XOSHIRO128 if prec == tp.PREC_F32 else tp.RNG_XOSHIRO
    print(f"device fold precision={prec} N={n}", flush=True)
    nb = (n + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    dev = torch.device("cuda:0")
    per_dev = torch.zeros(nb * 104, dtype=torch.uint8, device=dev)
    merged_dev = torch.zeros(104, dtype=torch.uint8, device=dev)
    params = tp.params_from_config(cfg.with_events(n), precision=prec)
    engine.run_device(params, kind, tp.SEED_SKIPAHEAD, tp.rng_state(kind), n, per_dev.data_ptr(),
                      merged_dev.data_ptr())
    engine.sync()
    per = np.frombuffer(per_dev.cpu().numpy().tobytes(), dtype=tp.ACC_DTYPE)
    dfold = tp.TpAcc.from_buffer_copy(merged_dev.cpu().numpy().tobytes())
    hfold = tp.merge(per, f32=prec == tp.PREC_F32)
    np.testing.assert_array_equal(dfold.as_array(), hfold.as_array())


@pytest.mark.parametrize("n_events", [1, 9999, 35017, 123457])
def test_f32_faster_evgen_jump_per_batch_matches_oracle(engine, cfg, oracle32, valeurs, n_events):
    """faster-evgen in f32 with faster-threading seeding (the reference's
    fastest build, VALIDATION.md:58-60): RANF values as f32, the unit-disc
    re-roll decided in f32, every batch reseeded; against the f32 oracle."""
    print(f"f32 faster-evgen jump N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32, faster_evgen=True)
    per, merged, _ = engine.run(params, tp.RNG_RANF, tp.SEED_JUMP, tp.rng_state(tp.RNG_RANF), n_events)
    ocfg = oracle32.config(valeurs, n_events)
    om, oper = oracle32.run(ocfg, oracle32.opts("ranf", faster_threading=True, threads=4, faster_evgen=True),
                            per_batch=True)
    worst = 0.0
    for b in range((n_events + 9999) // 10000):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        worst = max(worst, close_f32(a, oper[b].as_array(), min(10000, n_events - 10000 * b)))
    close_f32(merged.as_array(), om.as_array(), n_events)
    print(f"  worst relative difference of a batch sum: {worst:.2e}")


def test_f32_faster_evgen_jump_across_chunks(engine, cfg, oracle32, valeurs):
    n_events = FE_TWO_CHUNKS
    print(f"f32 faster-evgen jump N={n_events}", flush=True)
    params = tp.params_from_config(cfg.with_events(n_events), precision=tp.PREC_F32, faster_evgen=True)
    per, _, stats = engine.run(params, tp.RNG_RANF, tp.SEED_JUMP, tp.rng_state(tp.RNG_RANF), n_events)
    assert stats.chunks == 2
    ocfg = oracle32.config(valeurs, n_events)
    opts = oracle32.opts("ranf", faster_evgen=True)
    for b in FE_EDGE_BATCHES:
        g = oracle32.rng("ranf")
        for _ in range(b):
            oracle32.jump(g)
        o = oracle32.simulate_batch(ocfg, opts, g, 10000).as_array()
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        close_f32(a, o, 10000)


@pytest.mark.parametrize("n_events", [1, 9999, 35017, 123457])
def test_f32_faster_evgen_per_batch_matches_oracle(engine, cfg, oracle32, valeurs, n_events):
    """faster-evgen in f32 with skip-ahead seeding (the `f32,faster-evgen`
    build): one RANF stream walked with the unit-disc test in f32 (host walk
    at these sizes), batch by batch against the f32 oracle."""
    print(f"f32 faster-evgen N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32, faster_evgen=True)
    per, merged, _ = engine.run(params, tp.RNG_RANF, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_RANF), n_events)
    ocfg = oracle32.config(valeurs, n_events)
    om, oper = oracle32.run(ocfg, oracle32.opts("ranf", threads=4, faster_evgen=True), per_batch=True)
    for b in range((n_events + 9999) // 10000):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        close_f32(a, oper[b].as_array(), min(10000, n_events - 10000 * b))
    close_f32(merged.as_array(), om.as_array(), n_events)


def test_f32_faster_evgen_device_walk_across_chunks(engine, cfg, oracle32, valeurs):
    """N = 1.24e8: the device walk with the unit-disc test in f32, two chunks;
    every batch against the f32 oracle's run."""
    n_events = FE_TWO_CHUNKS
    print(f"f32 faster-evgen device walk N={n_events}", flush=True)
    params = tp.params_from_config(cfg.with_events(n_events), precision=tp.PREC_F32, faster_evgen=True)
    per, merged, stats = engine.run(params, tp.RNG_RANF, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_RANF), n_events)
    assert stats.host_walk_chunks == 0, "the device walkers did not all meet"
    assert stats.chunks == 2
    print("oracle run", flush=True)
    om, oper = oracle32.run(oracle32.config(valeurs, n_events), oracle32.opts("ranf", threads=16, faster_evgen=True),
                            per_batch=True)
    _, cond_batch = conditioning_state()
    assert close_f32_batches(per, oper) == [cond_batch]
    close_f32(merged.as_array(), om.as_array(), n_events)


def test_f32_faster_evgen_conditioning_event(engine, cfg, oracle, oracle32, valeurs):
    """The one event of the 1.24e8-event f32 faster-evgen RANF run whose sums
    the GPU's f32 and the f32 oracle put percents apart: from the same state,
    the GPU's one-event closure selects it as both oracles do, consumes the
    same draws, and lands no further from the f64 oracle than four times the
    f32 oracle does."""
    st, _ = conditioning_state()
    params = tp.params_from_config(cfg, precision=tp.PREC_F32, faster_evgen=True)
    a = engine.simulate_batch(params, tp.RNG_RANF, st, 1).as_array()
    errs = []
    for lib in (oracle32, oracle):
        g = lib.rng("ranf")
        st0, _ = conditioning_state()
        for i in range(56):
            g.numbers[i] = st0.numbers[i]
        g.index = st0.index
        o = lib.simulate_batch(lib.config(valeurs), lib.opts("ranf", faster_evgen=True), g, 1).as_array()
        assert a[0] == o[0] == 1 and st.index == g.index == 2
        errs.append(o)
    o32, o64 = errs
    gpu_err = float(np.max(np.abs(a[1:] - o64[1:]) / np.abs(o64[1:])))
    oracle_err = float(np.max(np.abs(o32[1:] - o64[1:]) / np.abs(o64[1:])))
    print(f"  relative distance to f64: GPU f32 {gpu_err:.3f}, f32 oracle {oracle_err:.3f}")
    assert gpu_err <= 4 * oracle_err, (gpu_err, oracle_err)


def test_f32_faster_evgen_simulate_batch_advances_like_the_closure(engine, cfg, oracle32, valeurs):
    params = tp.params_from_config(cfg, precision=tp.PREC_F32, faster_evgen=True)
    st = tp.rng_state(tp.RNG_RANF)
    g = oracle32.rng("ranf")
    ocfg = oracle32.config(valeurs)
    opts = oracle32.opts("ranf", faster_evgen=True)
    for n in (10000, 7, 1, 23456):
        acc = engine.simulate_batch(params, tp.RNG_RANF, st, n)
        oacc = oracle32.simulate_batch(ocfg, opts, g, n)
        close_f32(acc.as_array(), oacc.as_array(), n)
        assert st.index == g.index and list(st.numbers)[1:] == list(g.numbers)[1:]


@pytest.mark.parametrize("jump", [False, True])
@pytest.mark.parametrize("n_events", [1, 9999, 35017, 123457])
def test_f32_faster_evgen_xoshiro128_per_batch_matches_oracle(engine, cfg, oracle32, valeurs, n_events, jump):
    """faster-evgen on the xoshiro128+ stream of the all-f32 mode
    (standard-random,f32,faster-evgen, with or without faster-threading):
    random_array::<N> as N consecutive f32 draws, the unit-disc test in f32;
    batch by batch against the f32 oracle."""
    print(f"f32 faster-evgen xoshiro128 N={n_events} jump={jump}", flush=True)
    c = cfg.with_events(n_events)
    params = tp.params_from_config(c, precision=tp.PREC_F32, faster_evgen=True)
    seeding = tp.SEED_JUMP if jump else tp.SEED_SKIPAHEAD
    per, merged, _ = engine.run(params, tp.RNG_XOSHIRO128, seeding, tp.rng_state(tp.RNG_XOSHIRO128), n_events)
    ocfg = oracle32.config(valeurs, n_events)
    om, oper = oracle32.run(ocfg, oracle32.opts("xoshiro", faster_threading=jump, threads=4, faster_evgen=True),
                            per_batch=True)
    for b in range((n_events + 9999) // 10000):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        close_f32(a, oper[b].as_array(), min(10000, n_events - 10000 * b))
    close_f32(merged.as_array(), om.as_array(), n_events)


@pytest.mark.parametrize("jump", [False, True])
def test_f32_faster_evgen_xoshiro128_across_chunks(engine, cfg, oracle32, valeurs, jump):
    """N = 1.24e8 (two chunks): with skip-ahead seeding the device walk of the
    one stream, with jump seeding the per-batch walks; every batch against the
    f32 oracle's run."""
    n_events = FE_TWO_CHUNKS
    print(f"f32 faster-evgen xoshiro128 N={n_events} jump={jump}", flush=True)
    params = tp.params_from_config(cfg.with_events(n_events), precision=tp.PREC_F32, faster_evgen=True)
    seeding = tp.SEED_JUMP if jump else tp.SEED_SKIPAHEAD
    per, merged, stats = engine.run(params, tp.RNG_XOSHIRO128, seeding, tp.rng_state(tp.RNG_XOSHIRO128), n_events)
    assert stats.host_walk_chunks == 0, "the device walkers did not all meet"
    assert stats.chunks == 2
    print("oracle run", flush=True)
    om, oper = oracle32.run(oracle32.config(valeurs, n_events),
                            oracle32.opts("xoshiro", faster_threading=jump, threads=16, faster_evgen=True),
                            per_batch=True)
    close_f32_batches(per, oper)
    close_f32(merged.as_array(), om.as_array(), n_events)


def test_f32_faster_evgen_xoshiro128_simulate_batch_advances_like_the_closure(engine, cfg, oracle32, valeurs):
    params = tp.params_from_config(cfg, precision=tp.PREC_F32, faster_evgen=True)
    st = tp.rng_state(tp.RNG_XOSHIRO128)
    g = oracle32.rng("xoshiro")
    ocfg = oracle32.config(valeurs)
    opts = oracle32.opts("xoshiro", faster_evgen=True)
    for n in (10000, 7, 1, 23456):
        acc = engine.simulate_batch(params, tp.RNG_XOSHIRO128, st, n)
        oacc = oracle32.simulate_batch(ocfg, opts, g, n)
        close_f32(acc.as_array(), oacc.as_array(), n)
        assert list(st.s) == list(g.xs32)
