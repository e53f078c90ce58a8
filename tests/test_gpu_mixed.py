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
