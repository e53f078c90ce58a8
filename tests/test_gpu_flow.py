"""The RealNVP proposal on the GPU (tp_flow_*): the bf16-MFMA forward, inverse
and backward kernels against the PyTorch fp32 reference in flow_ref.py (bf16
rounding emulated), the density against the mathematics (normalisation over
the cube, change of variables), the training integrand against tp_adjoint
and finite differences, and the trained proposal against plain RAMBO (the
cross-section within two standard errors, a smaller variance per event)."""
import numpy as np
import pytest
import torch

import flow_ref as F
import trois_photons as tp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    print("creating engine on cuda:0", flush=True)
    with tp.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def setup(valeurs):
    cfg = tp.Configuration.load(valeurs)
    return cfg, tp.params_from_config(cfg)


@pytest.fixture(scope="module")
def rflow(engine):
    """A flow with random, non-trivial weights (output layer not zero)."""
    f = tp.Flow(engine, init_seed=7)
    w = F.random_params(11)
    f.set_params(w)
    yield f, w
    f.close()


def base_points(n, seed):
    rng = np.random.default_rng(seed)
    v = np.full((n, 16), 0.5, np.float32)
    v[:, :12] = rng.uniform(2e-3, 1 - 2e-3, (n, 12)).astype(np.float32)
    return v


def test_params_roundtrip_and_identity_init(engine):
    f = tp.Flow(engine, init_seed=3)
    w = f.get_params()
    assert w.shape == (tp.FLOW_NPARAM,)
    W3 = np.concatenate([w[l * F.LP + F.OFF_W3:(l + 1) * F.LP] for l in range(F.L)])
    assert not W3.any(), "output layers start at zero (identity flow)"
    v = base_points(4096, 1)
    u, z, lq = f.sample(4096, v=v)
    z0 = np.log(v[:, :12]) - np.log1p(-v[:, :12])
    assert np.abs(z - z0).max() < 1e-5
    assert np.abs(lq).max() < 1e-3, "identity flow has density 1"
    w2 = F.random_params(5)
    f.set_params(w2)
    np.testing.assert_array_equal(f.get_params(), w2)
    f.close()


def test_forward_matches_torch_reference(rflow):
    f, w = rflow
    n = 4096 + 37  # ragged: not a multiple of the 512-sample block
    v = base_points(n, 2)
    print(f"flow forward n={n}", flush=True)
    u, z, lq = f.sample(n, v=v)
    zr, lqr, _, _ = F.forward(torch.tensor(v[:, :12]), torch.tensor(w), emulate=True)
    dz = np.abs(z - zr.numpy())
    dl = np.abs(lq - lqr.numpy())
    print(f"|dz| median {np.median(dz):.2e} p99 {np.quantile(dz, .99):.2e} max {dz.max():.2e}; "
          f"|dlogq| median {np.median(dl):.2e} max {dl.max():.2e}")
    assert np.median(dz) < 1e-4 and np.quantile(dz, 0.99) < 5e-3 and dz.max() < 5e-2
    assert np.median(dl) < 1e-4 and dl.max() < 5e-2
    np.testing.assert_allclose(u, 1 / (1 + np.exp(-z.astype(np.float64))), rtol=1e-6)


def test_inverse_matches_forward_and_reference(rflow):
    f, w = rflow
    n = 8192
    v = base_points(n, 3)
    u, z, lq = f.sample(n, v=v)
    lqi = f.density(u)
    d = np.abs(lqi - lq)
    print(f"inverse vs forward |dlogq| median {np.median(d):.2e} p99 {np.quantile(d, .99):.2e} max {d.max():.2e}")
    assert np.median(d) < 1e-4 and np.quantile(d, 0.99) < 1e-2
    lqr = F.inverse_logq(torch.tensor(u[:1024]), torch.tensor(w, dtype=torch.float64), emulate=False).numpy()
    assert np.median(np.abs(lqi[:1024] - lqr)) < 1e-2


def test_density_is_normalised(rflow):
    """int q du = 1: mean of q over uniform points, and mean of 1/q over flow
    samples (both estimate the volume of the cube)."""
    f, _ = rflow
    n = 1 << 20
    rng = np.random.default_rng(4)
    u = rng.uniform(0, 1, (n, 12))
    q = np.exp(f.density(u).astype(np.float64))
    m1, se1 = q.mean(), q.std() / np.sqrt(n)
    _, _, lq = f.sample(n, key=99)
    iq = np.exp(-lq.astype(np.float64))
    m2, se2 = iq.mean(), iq.std() / np.sqrt(n)
    print(f"E_U[q] = {m1:.5f} +- {se1:.5f};  E_q[1/q] = {m2:.5f} +- {se2:.5f}")
    assert abs(m1 - 1) < 4 * se1 + 5e-3
    assert abs(m2 - 1) < 4 * se2 + 5e-3


def test_logdet_matches_jacobian(rflow):
    """log q from the kernels' formula against -log|det du/dv| of the f32 flow
    computed by autograd (a sign or a missing term would differ by O(1))."""
    f, w = rflow
    v = base_points(32, 5)
    _, _, lq = f.sample(32, v=v)
    wt = torch.tensor(w, dtype=torch.float64)
    for i in range(8):
        x = torch.tensor(v[i, :12], dtype=torch.float64)
        J = torch.autograd.functional.jacobian(
            lambda y: torch.sigmoid(F.forward(y[None], wt, emulate=False)[0])[0], x)
        ref = -float(torch.linalg.slogdet(J)[1])
        assert abs(ref - lq[i]) < 2e-2, (i, ref, lq[i])


def test_backward_matches_autograd(rflow):
    f, w = rflow
    n = 2048
    v = base_points(n, 6)
    rng = np.random.default_rng(7)
    gz = np.zeros((n, 16), np.float32)
    gz[:, :12] = rng.normal(0, 1, (n, 12))
    print(f"flow backward n={n}", flush=True)
    g, gv = f.grad(v, gz)
    gr, gvr = F.grad(v, gz, w, emulate=True)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    print(f"param grad rel L2 err {rel:.2e} cos {cos:.6f}")
    for l in range(F.L):
        sl = slice(l * F.LP, (l + 1) * F.LP)
        rl = np.linalg.norm(g[sl] - gr[sl]) / np.linalg.norm(gr[sl])
        assert rl < 2e-2, (l, rl)
    assert rel < 1e-2 and cos > 0.9999
    assert not gv[:, 12:].any()
    dv = np.abs(gv[:, :12] - gvr) / (np.abs(gvr).max(1, keepdims=True) + 1e-6)
    print(f"d/dz0 rel err median {np.median(dv):.2e} p99 {np.quantile(dv, .99):.2e}")
    assert np.median(dv) < 1e-2 and np.quantile(dv, 0.99) < 5e-2


@pytest.mark.parametrize("tau,target", [(0.0, 0), (0.02, 1), (0.02, 2)])
def test_integrand_matches_adjoint_and_fd(engine, setup, rflow, tau, target):
    """f against tp_adjoint (hard cuts) or below it by the smooth cut factor,
    and gz against central differences of the loss in z."""
    cfg, params = setup
    f, _ = rflow
    n = 4096
    rng = np.random.default_rng(8)
    u = rng.uniform(0.02, 0.98, (n, 12))
    eps = 1e-3
    fv, gz, lf = f.integrand(params, u, eps, tau, target)
    Fa, _, _ = engine.adjoint(params, n, u=u)
    sc = np.array(params.sigma_contribs)
    ref = (Fa @ sc) / params.ev_weight
    if tau == 0:
        np.testing.assert_allclose(fv, ref, rtol=1e-12, atol=1e-300)
    else:
        inside = ref > 0
        ratio = fv[inside] / ref[inside]
        print(f"soft/hard f inside the cuts: median {np.median(ratio):.4f} min {ratio.min():.3e}; "
              f"soft f > 0 outside: {np.count_nonzero(fv[~inside] > 0)} of {np.count_nonzero(~inside)}")
        assert np.all(ratio <= 1 + 1e-12) and np.median(ratio) > 0.9
        assert np.count_nonzero(fv[~inside] > 0) > 0
    assert np.all(gz[:, 12:] == 0)
    e = 0.0 if target == 2 else eps
    np.testing.assert_allclose(lf, np.log(fv + e), rtol=1e-6)

    def loss(uu):
        ff, _, _ = f.integrand(params, uu, eps, tau, target)
        return -np.log(uu * (1 - uu)).sum(1) - np.log(ff + e)

    keep = fv > 0
    zc = np.log(u) - np.log1p(-u)
    h = 1e-5
    worst = []
    for j in range(12):
        zp, zm = zc.copy(), zc.copy()
        zp[:, j] += h
        zm[:, j] -= h
        up, um = 1 / (1 + np.exp(-zp)), 1 / (1 + np.exp(-zm))
        lp, lm = loss(up), loss(um)
        fp, _, _ = f.integrand(params, up, eps, tau, target)
        fm, _, _ = f.integrand(params, um, eps, tau, target)
        ok = keep & (fp > 0) & (fm > 0)
        fd = (lp - lm) / (2 * h)
        err = np.abs(gz[ok, j] - fd[ok]) / np.maximum(1.0, np.abs(fd[ok]))
        worst.append(np.quantile(err, 0.99))
    print(f"gz vs FD: p99 rel err per feature max {max(worst):.2e}")
    assert max(worst) < 1e-3


def finalize(cfg, acc):
    return tp.finalize(cfg, acc)


def test_training_reduces_loss_and_flow_estimates_sigma(engine, setup):
    """Training (1000 steps of 65536, default options), then 2e7 weighted
    events against 2e7 plain events: sigma within 2 combined standard errors,
    and a smaller relative error at equal event count."""
    cfg, params = setup
    fl = tp.Flow(engine, init_seed=1)
    opts = tp.flow_default_opts(steps=1000)
    print("flow training 1000 x 65536", flush=True)
    loss, st = fl.train(params, opts)
    print(f"loss first {loss[:10].mean():.4f} last {loss[-10:].mean():.4f} ({st.device_ms:.1f} ms)")
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < loss[:50].mean() - 0.02

    n = 20_000_000
    c = cfg.with_events(n)
    pc = tp.params_from_config(c)
    print(f"weighted run n={n}", flush=True)
    per_f, accf, _ = fl.run(pc, 0.9, 12345, n)
    np.testing.assert_array_equal(tp.merge(per_f).as_array(), accf.as_array())  # device fold = host fold
    print(f"plain run n={n}", flush=True)
    _, accp, _ = engine.run(pc, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), n)
    rf, rp = finalize(c, accf), finalize(c, accp)
    se_f, se_p = rf.sigma * rf.prec, rp.sigma * rp.prec
    print(f"sigma flow {rf.sigma:.6f} +- {se_f:.6f} (prec {rf.prec:.3e}); "
          f"plain {rp.sigma:.6f} +- {se_p:.6f} (prec {rp.prec:.3e}); variance ratio {(rp.prec / rf.prec) ** 2:.2f}")
    assert abs(rf.sigma - rp.sigma) < 2 * np.hypot(se_f, se_p)
    assert rf.prec < rp.prec
    fl.close()
