"""The oracle's f64 gradient (dual numbers over its own event code) against
central finite differences of the oracle's forward value -- so that the
gradient used to check the GPU adjoint is itself pinned."""
import numpy as np
import pytest

from oracle_ffi import OracleGrad


@pytest.fixture(scope="module")
def setup(oracle, valeurs):
    cfg = oracle.config(valeurs)
    g = oracle.rng("xoshiro", 42)
    u = oracle.uniforms(g, 12 * 400).reshape(400, 12)
    return cfg, OracleGrad(oracle, cfg), u


def forward(oracle, cfg, grad, u):
    _, keep, _, m, _ = oracle.events(cfg, oracle.opts("xoshiro"), u)
    return grad.k[8] * m * keep[:, None], keep


def test_du_matches_finite_differences(oracle, setup):
    cfg, grad, u = setup
    F, gu, _ = grad.grad(u)
    keep = np.any(F != 0, axis=1)
    checked = 0
    for j in range(12):
        h = 1e-6 * np.maximum(u[:, j], 1e-3)
        up, um = u.copy(), u.copy()
        up[:, j] += h
        um[:, j] -= h
        Fp, kp = forward(oracle, cfg, grad, up)
        Fm, km = forward(oracle, cfg, grad, um)
        ok = keep & kp & km
        fd = (Fp - Fm) / (2 * h[:, None])
        scale = np.abs(gu[ok]).max(axis=2)  # per point and channel
        err = np.abs(fd[ok] - gu[ok, :, j])
        assert np.all(err <= 1e-5 * scale + 1e-12 * np.abs(F[ok])), (j, (err / scale).max())
        checked += ok.sum()
    assert checked > 2000


def test_dp_matches_finite_differences(oracle, setup):
    cfg, grad, u = setup
    F, _, gp = grad.grad(u)
    keep = np.any(F != 0, axis=1)
    p, _, _, _, _ = oracle.events(cfg, oracle.opts("xoshiro"), u)
    p = p[keep]
    g = gp[keep]
    for j in range(12):
        i, k = 2 + j // 4, j % 4
        # the spinor products use sqrt(E + Z): near -Z the step must stay
        # far below E + Z
        h = np.minimum(1e-5 * np.abs(p[:, i, 3]), 1e-3 * (p[:, i, 3] + p[:, i, 2]))

        def at(step):
            q = p.copy()
            q[:, i, k] += step * h
            return grad.m2_at(q)
        # 4th-order central stencil: truncation O(h^4)
        fd = (-at(2) + 8 * at(1) - 8 * at(-1) + at(-2)) / (12 * h[:, None])
        scale = np.abs(g).max(axis=2)
        err = np.abs(fd - g[:, :, j])
        assert np.all(err <= 1e-6 * scale), (j, (err / scale).max())


def test_photon_permutation_invariance(oracle, setup):
    """sum over helicities |M|^2 is symmetric in the photons, so sorting or
    not gives the same value (to rounding) and permuted momentum gradients."""
    cfg, grad, u = setup
    F1, _, _ = grad.grad(u)
    g2 = OracleGrad(oracle, cfg, sort=False)
    F2, _, _ = g2.grad(u)
    both = np.all(F1 != 0, axis=1) & np.all(F2 != 0, axis=1)
    rel = np.abs(F1[both, :3] - F2[both, :3]) / np.abs(F1[both, :3])
    assert rel.max() < 1e-12
