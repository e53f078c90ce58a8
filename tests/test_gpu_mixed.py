"""Mixed precision (f32 spinor/amplitude algebra, f64 conformal map, cuts and
weight, Kahan-compensated f64 batch sums) against the f64 path on identical
streams: per-event and per-quantity relative errors."""
import numpy as np
import pytest

import trois_photons as tp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    print("creating engine on cuda:0", flush=True)
    with tp.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def cfg(valeurs):
    return tp.Configuration.load(valeurs)


@pytest.mark.parametrize("rng", ["xoshiro", "ranf"])
def test_events_mixed_vs_f64(engine, cfg, rng):
    kind = {"ranf": tp.RNG_RANF, "xoshiro": tp.RNG_XOSHIRO}[rng]
    n = 50000
    print(f"mixed events rng={rng} n={n}", flush=True)
    st = tp.rng_state(kind)
    u64, p64, k64, m64 = engine.events(tp.params_from_config(cfg), kind, st, n)
    u32, p32, k32, m32 = engine.events(tp.params_from_config(cfg, precision=tp.PREC_MIXED), kind, st, n)
    assert np.array_equal(u64, u32)
    assert np.array_equal(p64, p32), "RAMBO and the cuts stay f64 in the mixed path"
    flips = int(np.sum(k64 != k32))
    both = (k64 == 1) & (k32 == 1)
    rel = np.abs(m32[both, :3] - m64[both, :3]) / np.abs(m64[both, :3])
    scale = 2 * np.sqrt(m64[both, 0] * m64[both, 1])
    relx = np.abs(m32[both, 3:] - m64[both, 3:]) / scale[:, None]
    print(f"cut flips {flips}/{n}; |M|^2 rel err median {np.median(rel):.2e} p99 {np.percentile(rel, 99):.2e} "
          f"max {rel.max():.2e}; interference (vs 2|A||B+|) median {np.median(relx):.2e} max {relx.max():.2e}")
    assert flips == 0
    assert np.median(rel) < 2e-6 and np.percentile(rel, 99) < 5e-5
    assert np.median(relx) < 2e-6 and np.percentile(relx, 99) < 5e-5


QUANTITIES = ["sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m"]


def test_run_mixed_vs_f64(engine, cfg):
    n = 20_000_000
    c = cfg.with_events(n)
    res = {}
    for prec in (tp.PREC_F64, tp.PREC_MIXED):
        print(f"mixed run precision={prec} N={n}", flush=True)
        st = tp.rng_state(tp.RNG_XOSHIRO)
        per, merged, _ = engine.run(tp.params_from_config(c, precision=prec), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, st, n)
        res[prec] = (merged, tp.finalize(c, merged))
    (m64, f64), (m32, f32) = res[tp.PREC_F64], res[tp.PREC_MIXED]
    print(f"selected f64 {m64.selected_events} mixed {m32.selected_events}")
    assert m64.selected_events == m32.selected_events
    errs = {q: abs(getattr(f32, q) - getattr(f64, q)) / abs(getattr(f64, q)) for q in QUANTITIES}
    names = ["A", "B+", "B-", "R_MX", "I_MX"]
    for k in range(5):
        errs[f"spm2[{names[k]}]"] = abs(m32.spm2[k] - m64.spm2[k]) / abs(m64.spm2[k])
        errs[f"vars[{names[k]}]"] = abs(m32.vars[k] - m64.vars[k]) / abs(m64.vars[k])
    for k, v in errs.items():
        print(f"  rel err {k:12s} {v:.3e}")
    # sigma and the |M|^2 sums agree to ~1e-7; the cancelling interference
    # sum is compared against its own statistical error instead.
    assert errs["sigma"] < 1e-6
    for k in ("A", "B+", "B-"):
        assert errs[f"spm2[{k}]"] < 2e-6 and errs[f"vars[{k}]"] < 5e-5
    # statistical error of the I_MX sum: sqrt(sum x^2); the f32 shift must be
    # a negligible fraction of it
    stat = np.sqrt(m64.vars[4])
    assert abs(m32.spm2[4] - m64.spm2[4]) < 1e-3 * stat


# ---- against the FP64 oracle (config 3: every mixed result carries its error
# against the oracle's f64 result on the same stream) -------------------------
NAMES = ["A", "B+", "B-", "R_MX", "I_MX"]


def mixed_errors(a, b, n_events):
    """Per-quantity relative errors of accumulator a (mixed) against b (f64
    oracle): the |M|^2 sums A, B+, B- and sigma relative to themselves, the
    cancelling interference sums R_MX, I_MX relative to their statistical
    scale sqrt(n sum x^2), the second moments relative to themselves."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    e = {}
    for k in range(5):
        if k < 3:
            e[f"spm2[{NAMES[k]}]"] = abs(a[1 + k] - b[1 + k]) / abs(b[1 + k])
        else:
            e[f"spm2[{NAMES[k]}]/stat"] = abs(a[1 + k] - b[1 + k]) / np.sqrt(n_events * b[6 + k])
        e[f"vars[{NAMES[k]}]"] = abs(a[6 + k] - b[6 + k]) / abs(b[6 + k])
    e["sigma"] = abs(a[11] - b[11]) / abs(b[11])
    e["variance"] = abs(a[12] - b[12]) / abs(b[12])
    return e


# Budget per quantity for a batch (or run) of >= 1e4 events: the f32 spinor
# and amplitude algebra leaves per-event |M|^2 errors of ~2e-7 (median), which
# average down in the sums (measured, batch or run: |M|^2 sums <= 2.1e-7,
# second moments <= 4.6e-7, sigma <= 7.1e-8, interference sums <= 1.0e-7 of
# their statistical scale).
BUDGET = {"spm2": 1e-6, "stat": 1e-6, "vars": 2e-6, "sigma": 5e-7, "variance": 1e-6}


def assert_budget(e):
    for k, v in e.items():
        lim = BUDGET["stat"] if k.endswith("/stat") else BUDGET[k.split("[")[0]]
        assert v <= lim, (k, v, lim)


@pytest.mark.parametrize("rng", ["xoshiro", "ranf"])
@pytest.mark.parametrize("n_events", [9999, 35017])
def test_mixed_per_batch_vs_oracle(engine, cfg, oracle, valeurs, rng, n_events):
    kind = {"ranf": tp.RNG_RANF, "xoshiro": tp.RNG_XOSHIRO}[rng]
    print(f"mixed vs oracle per batch rng={rng} N={n_events}", flush=True)
    c = cfg.with_events(n_events)
    per, merged, _ = engine.run(tp.params_from_config(c, precision=tp.PREC_MIXED), kind, tp.SEED_SKIPAHEAD,
                                tp.rng_state(kind), n_events)
    om, oper = oracle.run(oracle.config(valeurs, n_events), oracle.opts(rng, threads=4), per_batch=True)
    worst = {}
    for b in range((n_events + 9999) // 10000):
        row = per[b]
        a = np.array([row["selected_events"], *row["spm2"], *row["vars"], row["sigma"], row["variance"]])
        o = oper[b].as_array()
        assert a[0] == o[0]
        e = mixed_errors(a, o, min(10000, n_events - 10000 * b))
        assert_budget(e)
        worst = {k: max(v, worst.get(k, 0.0)) for k, v in e.items()}
    print("  worst per-batch errors: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))


def test_mixed_finals_vs_oracle(engine, cfg, oracle_run):
    """N = 1e7 (xoshiro256+): every res.data quantity of the mixed run against
    the oracle's f64 run."""
    n = 10_000_000
    print(f"mixed finals vs oracle N={n}", flush=True)
    c = cfg.with_events(n)
    per, merged, _ = engine.run(tp.params_from_config(c, precision=tp.PREC_MIXED), tp.RNG_XOSHIRO,
                                tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), n)
    ocfg, omerged, _, ofin = oracle_run("xoshiro", False, True, False)
    assert merged.selected_events == omerged.selected
    e = mixed_errors(merged.as_array(), omerged.as_array(), n)
    assert_budget(e)
    fin = tp.finalize(c, merged)
    for q in QUANTITIES:
        e[q] = abs(getattr(fin, q) - getattr(ofin, q)) / abs(getattr(ofin, q))
    print("  " + ", ".join(f"{k} {v:.1e}" for k, v in e.items()))
    for q in ("sigma", "prec", "variance", "beta_min", "ss_p", "ss_m"):
        assert e[q] <= 1e-6, (q, e[q])
    for q in ("inc_ss_p", "inc_ss_m"):
        assert e[q] <= 1e-5, (q, e[q])


def test_mixed_full_size_1e10(engine, cfg):
    """Config 3 at full size: 1e10 events (1e6 batches) of the xoshiro256+
    stream, mixed against f64 on the GPU (the oracle would need hours):
    identical selection, sigma within 1e-6, no non-finite batch."""
    n = 10_000_000_000
    c = cfg.with_events(n)
    out = {}
    for prec in (tp.PREC_MIXED, tp.PREC_F64):
        print(f"full-size run precision={prec} N={n}", flush=True)
        per, merged, _ = engine.run(tp.params_from_config(c, precision=prec), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD,
                                    tp.rng_state(tp.RNG_XOSHIRO), n)
        finite = (np.isfinite(per["spm2"]).all() and np.isfinite(per["vars"]).all()
                  and np.isfinite(per["sigma"]).all() and np.isfinite(per["variance"]).all())
        assert finite, f"non-finite batch at precision {prec}"
        assert int(per["selected_events"].sum()) == merged.selected_events
        out[prec] = (merged, tp.finalize(c, merged))
    (mm, fm), (m64, f64) = out[tp.PREC_MIXED], out[tp.PREC_F64]
    print(f"  selected {mm.selected_events} (f64 {m64.selected_events}); sigma rel {abs(fm.sigma / f64.sigma - 1):.1e}")
    assert mm.selected_events == m64.selected_events
    assert abs(fm.sigma / f64.sigma - 1) <= 1e-6
