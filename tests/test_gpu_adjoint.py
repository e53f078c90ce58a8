"""The hand-written GPU adjoint (tp_adjoint) against the oracle's f64
forward-mode gradient, against central finite differences of the GPU's own
forward value, and the f32 adjoint against the f64 one."""
import numpy as np
import pytest

import trois_photons as tp
from oracle_ffi import OracleGrad

pytestmark = pytest.mark.gpu

N = 10000


@pytest.fixture(scope="module")
def engine():
    print("creating engine on cuda:0", flush=True)
    with tp.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def setup(oracle, valeurs):
    cfg = tp.Configuration.load(valeurs)
    ocfg = oracle.config(valeurs)
    g = oracle.rng("xoshiro", 42)
    u = oracle.uniforms(g, 12 * N).reshape(N, 12)
    return cfg, tp.params_from_config(cfg), OracleGrad(oracle, ocfg), u


def rel_to_scale(a, b):
    """|a - b| per (point, channel, component) over the channel's gradient scale."""
    scale = np.abs(b).max(axis=-1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(scale > 0, np.abs(a - b) / scale, 0.0)


def test_adjoint_f64_matches_oracle_gradient(engine, setup):
    cfg, params, og, u = setup
    print(f"adjoint f64 n={N} (stream xoshiro seed 42)", flush=True)
    st = tp.rng_state(tp.RNG_XOSHIRO, 42)
    F, gu, gp = engine.adjoint(params, N, tp.PREC_F64, st=st)
    oF, ogu, ogp = og.grad(u)
    keep, okeep = np.any(F != 0, axis=1), np.any(oF != 0, axis=1)
    assert np.array_equal(keep, okeep)
    relF = np.abs(F[keep] - oF[keep]) / np.abs(oF[keep]).max(axis=1, keepdims=True)
    ru, rp = rel_to_scale(gu[keep], ogu[keep]), rel_to_scale(gp[keep], ogp[keep])
    print(f"F rel {relF.max():.2e}; du rel median {np.median(ru):.2e} p99 {np.percentile(ru, 99):.2e} "
          f"max {ru.max():.2e}; dp rel median {np.median(rp):.2e} max {rp.max():.2e}")
    assert relF.max() < 1e-10
    assert np.percentile(ru, 99) < 1e-10 and ru.max() < 1e-6
    assert np.percentile(rp, 99) < 1e-10 and rp.max() < 1e-6
    assert np.all(gu[~keep] == 0) and np.all(gp[~keep] == 0)


def test_adjoint_matches_gpu_finite_differences(engine, setup):
    """Central differences (h = 1e-6 relative) of the GPU's own forward value,
    skipping points whose cut decision changes inside the stencil."""
    cfg, params, _, u = setup
    m = 400
    uu = u[:m]
    F, gu, _ = engine.adjoint(params, m, tp.PREC_F64, u=uu)
    keep = np.any(F != 0, axis=1)
    worst = 0.0
    for j in range(12):
        h = 1e-6 * np.maximum(uu[:, j], 1e-3)
        up, um = uu.copy(), uu.copy()
        up[:, j] += h
        um[:, j] -= h
        Fp, _, _ = engine.adjoint(params, m, tp.PREC_F64, u=up)
        Fm, _, _ = engine.adjoint(params, m, tp.PREC_F64, u=um)
        ok = keep & np.any(Fp != 0, axis=1) & np.any(Fm != 0, axis=1)
        fd = (Fp - Fm) / (2 * h[:, None])
        scale = np.abs(gu[ok]).max(axis=2)
        err = np.abs(fd[ok] - gu[ok, :, j]) / scale
        worst = max(worst, err.max())
    print(f"finite differences vs adjoint: worst rel {worst:.2e}")
    assert worst < 1e-5


def test_adjoint_f32_vs_f64(engine, setup):
    cfg, params, _, u = setup
    st = tp.rng_state(tp.RNG_XOSHIRO, 42)
    F64, gu64, gp64 = engine.adjoint(params, N, tp.PREC_F64, st=st)
    F32, gu32, gp32 = engine.adjoint(params, N, tp.PREC_F32, st=st)
    both = np.any(F64 != 0, axis=1) & np.any(F32 != 0, axis=1)
    flips = int(np.sum(np.any(F64 != 0, axis=1) != np.any(F32 != 0, axis=1)))
    rF = np.abs(F32[both] - F64[both]) / np.abs(F64[both]).max(axis=1, keepdims=True)
    ru = rel_to_scale(gu32[both], gu64[both])
    rp = rel_to_scale(gp32[both], gp64[both])
    print(f"f32 vs f64: cut flips {flips}; F rel median {np.median(rF):.2e} p99 {np.percentile(rF, 99):.2e}; "
          f"du rel median {np.median(ru):.2e} p99 {np.percentile(ru, 99):.2e}; dp median {np.median(rp):.2e}")
    assert flips <= 3
    assert np.median(rF) < 1e-5 and np.median(ru) < 1e-4 and np.median(rp) < 1e-4


def fd_gradient(og, uu, h_rel=1e-6):
    """Central differences (h = 1e-6 relative) of the oracle's f64 forward
    value F(u) with respect to each uniform; ok marks the points whose cut
    decision is the same at u, u+h and u-h."""
    F, _, _ = og.grad(uu)
    keep = np.any(F != 0, axis=1)
    fd = np.zeros((len(uu), 5, 12))
    ok = np.ones((len(uu), 12), bool)
    for j in range(12):
        h = h_rel * np.maximum(uu[:, j], 1e-3)
        up, um = uu.copy(), uu.copy()
        up[:, j] += h
        um[:, j] -= h
        Fp, _, _ = og.grad(up)
        Fm, _, _ = og.grad(um)
        ok[:, j] = keep & np.any(Fp != 0, axis=1) & np.any(Fm != 0, axis=1)
        fd[:, :, j] = (Fp - Fm) / (2 * h[:, None])
    return fd, ok


def test_adjoint_f32_matches_oracle_gradient(engine, setup):
    """The all-f32 adjoint against the oracle's f64 (dual-number) gradient on
    the 10^4 seed-42 points.  f32 rounding of the momenta flips the cut
    decision of a few points next to a cut boundary; elsewhere F and both
    gradients agree to f32 accuracy, amplified where RAMBO's conformal map
    and the spinor square roots are ill-conditioned."""
    cfg, params, og, u = setup
    print(f"adjoint f32 vs oracle f64 gradient n={N}", flush=True)
    st = tp.rng_state(tp.RNG_XOSHIRO, 42)
    F, gu, gp = engine.adjoint(params, N, tp.PREC_F32, st=st)
    oF, ogu, ogp = og.grad(u)
    keep, okeep = np.any(F != 0, axis=1), np.any(oF != 0, axis=1)
    flips = int(np.sum(keep != okeep))
    both = keep & okeep
    rF = np.abs(F[both] - oF[both]) / np.abs(oF[both]).max(axis=1, keepdims=True)
    ru, rp = rel_to_scale(gu[both], ogu[both]), rel_to_scale(gp[both], ogp[both])
    print(f"cut flips {flips}; F rel median {np.median(rF):.1e} p99 {np.percentile(rF, 99):.1e}; "
          f"du median {np.median(ru):.1e} p99 {np.percentile(ru, 99):.1e}; "
          f"dp median {np.median(rp):.1e} p99 {np.percentile(rp, 99):.1e}")
    assert flips <= 3
    assert np.median(rF) < 1e-5 and np.percentile(rF, 99) < 1e-3
    assert np.median(ru) < 1e-4 and np.percentile(ru, 99) < 1e-2
    assert np.median(rp) < 1e-4 and np.percentile(rp, 99) < 1e-2


def test_adjoint_f64_and_f32_match_central_differences(engine, setup):
    """Both adjoints against central differences (h = 1e-6 relative) of the
    oracle's f64 forward value on 400 points (f32 cannot resolve a step of
    1e-6, so the differences are taken in f64 for both)."""
    cfg, params, og, u = setup
    m = 400
    uu = u[:m]
    fd, ok = fd_gradient(og, uu)
    for prec, bar_med, bar_p99 in ((tp.PREC_F64, 1e-8, 1e-6), (tp.PREC_F32, 1e-4, 1e-2)):
        F, gu, _ = engine.adjoint(params, m, prec, u=uu)
        keep = np.any(F != 0, axis=1)
        errs = []
        for j in range(12):
            sel = ok[:, j] & keep
            scale = np.abs(fd[sel]).max(axis=2)
            with np.errstate(invalid="ignore", divide="ignore"):
                errs.append(np.where(scale > 0, np.abs(gu[sel, :, j] - fd[sel, :, j]) / scale, 0.0).ravel())
        e = np.concatenate(errs)
        print(f"precision {prec}: adjoint vs central differences median {np.median(e):.1e} "
              f"p99 {np.percentile(e, 99):.1e} max {e.max():.1e}")
        assert np.median(e) < bar_med and np.percentile(e, 99) < bar_p99
