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
