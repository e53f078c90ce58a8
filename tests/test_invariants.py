"""What the mathematics fixes, independently of any reference run.

* RAMBO (evgen.rs:89-132): the three photons are massless and balance the
  beams, sum p = (0, 0, 0, sqrt s) (beam along x, photons in the centre-of-mass
  frame), for every event, cut or not.
* Spinor products (spinor.rs:31-54): for massless momenta |s_ij|^2 = 2 p_i.p_j
  and s_ij = -s_ji.
* QED amplitudes (spinor.rs:72-162, matelems.rs:53-86): the helicity sum of
  |A g_a|^2 is the Berends-Kleiss closed form of e+e- -> 3 photons,
      2 e^6 s sum_i (p+.k_i)(p-.k_i)[(p+.k_i)^2 + (p-.k_i)^2] / prod_j (p+.k_j)(p-.k_j),
  event by event (the factor 2 is the reference's normalisation of one beam
  helicity pair); a dropped term, a wrong photon index or a sign error breaks
  the equality.
* Anomalous-coupling cross-sections against the reference's own analytic
  formula, fawzi() (resfin.rs:101-194).  Its cut cross-sections
  sigma+ = sig (ff + 2 gg), sigma- = sigma+ + 2 sig gg cover the energy,
  photon-photon and plane cuts but not the beam cut, so they are compared
  with runs without it (cut_beam = 1).

The oracle is checked here on the CPU; the GPU's momenta and helicity sums
(tp_events) under the same identities in the tests marked gpu.
"""
import math

import numpy as np
import pytest

import trois_photons as tp

EPS = np.finfo(np.float64).eps


def mdot(a, b):
    return a[..., 3] * b[..., 3] - a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2]


def conformal_condition(u):
    """kappa = R_E^2 / M^2 of RAMBO's conformal map (evgen.rs:94-98) for
    uniforms u[n][12]: M^2 = R_E^2 - |R|^2 cancels, so the map's rounding
    errors grow like eps kappa (the massless vectors q from evgen.rs:193-206,
    recomputed here in numpy only to size the tolerance)."""
    c = 2 * u[:, 0::4] - 1
    phi = 2 * np.pi * u[:, 1::4]
    e = -np.log(u[:, 2::4] * u[:, 3::4] + np.finfo(np.float64).tiny)
    st = np.sqrt(1 - c * c)
    q = np.stack([e * st * np.sin(phi), e * st * np.cos(phi), e * c, e], -1)
    r = q.sum(1)
    return r[:, 3] ** 2 / (r[:, 3] ** 2 - (r[:, :3] ** 2).sum(1))


def check_rambo(p, u, sqrt_s):
    """p[n][3][4] photon momenta of uniforms u[n][12]: conservation and
    masslessness to a few eps kappa (measured: at most 2.1 and 10.7 eps kappa
    over 2e4 events; a mistake in the map is an O(1) error)."""
    kappa = conformal_condition(u)
    tot = p.sum(axis=1)
    err = np.abs(tot - np.array([0.0, 0.0, 0.0, sqrt_s])).max(axis=1) / sqrt_s
    E = p[..., 3]
    m2 = (np.abs(mdot(p, p)) / (E * E)).max(axis=1)
    print(f"RAMBO: max |sum p - (0,0,0,sqrt s)| / (sqrt s eps kappa) = {(err / (EPS * kappa)).max():.2f}, "
          f"max |p^2| / (E^2 eps kappa) = {(m2 / (EPS * kappa)).max():.2f}, max kappa {kappa.max():.3g}")
    assert np.all(err <= 8 * EPS * kappa)
    assert np.all(m2 <= 32 * EPS * kappa)
    assert np.all(E > 0)


def berends_kleiss(p, alpha):
    """2 e^6 s sum_i (p+.k_i)(p-.k_i)[(p+.k_i)^2 + (p-.k_i)^2] / prod_j (p+.k_j)(p-.k_j);
    p[n][5][4] with e- = 0, e+ = 1, photons 2..4."""
    pm, pp, k = p[:, 0], p[:, 1], p[:, 2:]
    a = np.stack([mdot(pp, k[:, i]) for i in range(3)], 1)
    b = np.stack([mdot(pm, k[:, i]) for i in range(3)], 1)
    s = mdot(pm + pp, pm + pp)
    e6 = (4 * p.dtype.type(math.pi) * alpha) ** 3
    return 2 * e6 * s * (a * b * (a * a + b * b)).sum(1) / (a * b).prod(1)


def fawzi_sigma(c):
    """Cut cross-sections sigma+, sigma- (pb) of resfin.rs:101-194 (fawzi),
    from the oracle's or the package's configuration record."""
    br = c.br_ep_em if hasattr(c, "br_ep_em") else c.branching_ep_em
    conv = c.conv if hasattr(c, "conv") else c.gev2_to_picobarn
    mre = c.m_z0 / c.e_total
    gre = c.g_z0 * c.m_z0 / c.e_total ** 2
    x = 1 - mre ** 2
    sdz2 = abs(complex(x, -gre) / (x * x + gre * gre)) ** 2
    dl = (1 - c.cut_pp) / 2
    e4 = (2 * c.e_min / c.e_total) ** 4
    bra = c.m_z0 / (3 * 6 * math.pi ** 3 * 16 * 120)
    sig = (12 * math.pi / c.m_z0 ** 2 * br * c.g_z0 * bra / c.e_total ** 2 * (c.e_total / c.m_z0) ** 8
           * sdz2 * conv)
    d2, d3 = dl * dl, dl ** 3
    f1 = 1 - 15 * e4 - 9 / 7 * (1 - 70 * e4) * d2 + 6 / 7 * (1 + 70 * e4) * d3
    g1 = 1 - 30 * e4 - 9 / 7 * (1 - 70 * e4) * dl - 90 * e4 * d2 - 1 / 7 * (1 - 420 * e4) * d3
    g2 = 1 - 25 * e4 - 6 / 7 * (1 - 70 * e4) * dl - 3 / 7 * (1 + 210 * e4) * d2 - 8 / 21 * (1 - 52.5 * e4) * d3
    g3 = (1 - 195 / 11 * e4 - 18 / 77 * (1 - 7 * e4) * dl - 9 / 11 * (9 / 7 - 70 * e4) * d2
          - 8 / 11 * (1 - 105 / 11 * e4) * d3)
    c3 = c.cut_plane ** 3
    ff = f1 * (1 - c3)
    gg = g1 - 27 / 16 * g2 * c.cut_plane + 11 / 16 * g3 * c3
    sp = sig * (ff + 2 * gg)
    return sp, sp + 2 * sig * gg


def mc_sigma_pm(spm2, vars_):
    """sigma+-, as fawzi() forms them from the final table, and their
    relative statistical errors."""
    sp, va = np.asarray(spm2), np.asarray(vars_)
    out = []
    for col in (1, 2):
        mc = sp[:, col].sum() / 4
        inc = np.linalg.norm(sp[:, col] * va[:, col]) / abs(sp[:, col].sum())
        out.append((mc, inc))
    return out


@pytest.fixture(scope="module")
def events(oracle, valeurs):
    cfg = oracle.config(valeurs)
    g = oracle.rng("xoshiro")
    u = oracle.uniforms(g, 12 * 20000).reshape(-1, 12)
    p, keep, m2, m, w = oracle.events(cfg, oracle.opts("xoshiro"), u)
    return cfg, u, p, keep, m


def test_oracle_rambo_conserves_momentum(events):
    cfg, u, p, _, _ = events
    check_rambo(p[:, 2:], u, cfg.e_total)
    # incoming beams along x (evgen.rs:66-69)
    assert np.all(p[:, 0] == [-cfg.e_total / 2, 0, 0, cfg.e_total / 2])
    assert np.all(p[:, 1] == [cfg.e_total / 2, 0, 0, cfg.e_total / 2])


def test_oracle_spinor_identities(oracle, events):
    """|s_ij|^2 = 2 p_i.p_j for massless momenta.  RAMBO's photons are massless
    only to a few eps kappa, so the identity is taken in the form that holds
    for any momenta with E + Z > 0,
        |s_ij|^2 = 2 p_i.p_j - m_i^2 (E_j+Z_j)/(E_i+Z_i) - m_j^2 (E_i+Z_i)/(E_j+Z_j),
    m^2 = E^2 - |p|^2, its right side evaluated in long double from the same
    momenta.  What is left is the spinors' own rounding: a few eps times
    E_i E_j (E_i/(E_i+Z_i) + E_j/(E_j+Z_j)), the condition of
    sqrt(E + Z) (measured: at most 5.1 such units)."""
    _, _, p, _, _ = events
    s = oracle.spinors(p)
    L = np.longdouble
    P = p.astype(L)
    E, X, Y, Z = P[..., 3], P[..., 0], P[..., 1], P[..., 2]
    msq = E * E - X * X - Y * Y - Z * Z
    cond = (E / (E + Z)).astype(np.float64)
    worst = 0.0
    for i in range(5):
        for j in range(5):
            if i == j:
                assert np.all(s[:, i, i] == 0)
                continue
            dot = E[:, i] * E[:, j] - X[:, i] * X[:, j] - Y[:, i] * Y[:, j] - Z[:, i] * Z[:, j]
            rhs = (2 * dot - msq[:, i] * (E[:, j] + Z[:, j]) / (E[:, i] + Z[:, i])
                   - msq[:, j] * (E[:, i] + Z[:, i]) / (E[:, j] + Z[:, j]))
            lhs = s[:, i, j].real.astype(L) ** 2 + s[:, i, j].imag.astype(L) ** 2
            err = (np.abs(lhs - rhs) / (E[:, i] * E[:, j])).astype(np.float64)
            worst = max(worst, float((err / (EPS * (cond[:, i] + cond[:, j]))).max()))
    print(f"spinors: max ||s_ij|^2 - rhs| / (E_i E_j eps (c_i + c_j)) = {worst:.2f}")
    assert worst < 16
    assert np.array_equal(s, -np.transpose(s, (0, 2, 1)))  # exactly antisymmetric


def test_oracle_qed_sum_is_berends_kleiss(events, valeurs):
    cfg, _, p, keep, m = events
    check_qed(p, m[:, 0], cfg.alpha)


def check_qed(p, m_a, alpha):
    """Helicity sum of |A g_a|^2 against the closed form, evaluated in long
    double.  Most events agree to a few eps; the tail comes from RAMBO's
    photons being massless only to eps kappa (kappa up to ~1e3), which the
    closed form (exact for massless momenta) turns into relative differences
    up to ~3e-10 where a photon is nearly collinear with a beam.  A wrong term
    is an O(1) difference on most events."""
    r = np.abs((m_a.astype(np.longdouble) / berends_kleiss(p.astype(np.longdouble), alpha) - 1).astype(np.float64))
    print(f"QED: |helicity sum / Berends-Kleiss - 1|: median {np.median(r):.1e} p99 {np.percentile(r, 99):.1e} "
          f"max {r.max():.1e}")
    assert np.median(r) < 1e-14 and np.percentile(r, 99) < 1e-12 and r.max() < 2e-9


@pytest.mark.parametrize("cut_pp,e_min,cut_plane", [(0.9396, 4.559, 0.0), (1.0, 0.0, 0.0), (0.9396, 4.559, 0.3)])
def test_oracle_anomalous_sigma_against_fawzi(oracle, valeurs, cut_pp, e_min, cut_plane):
    """4e6 events without the beam cut: sigma+- from the MC against fawzi()'s
    analytic cut cross-sections.  They agree to 0.13-0.14 % with no cut at all, the
    difference being the same for sigma+ and sigma- (the analytic total
    normalises with BR(Z->ee) Gamma_Z, the MC with sin^2 theta_W and alpha_Z,
    which the shipped valeurs do not make exactly consistent); the energy and
    plane cuts change the ratio by a further 3e-4 at most."""
    cfg = oracle.config(valeurs, 4_000_000)
    cfg.cut_beam, cfg.cut_pp, cfg.e_min, cfg.cut_plane = 1.0, cut_pp, e_min, cut_plane
    merged, _ = oracle.run(cfg, oracle.opts("xoshiro", threads=8))
    fin = oracle.finalize(cfg, merged)
    (mcp, incp), (mcm, incm) = mc_sigma_pm(fin.spm2, fin.vars)
    sp, sm = fawzi_sigma(cfg)
    print(f"cuts pp={cut_pp} emin={e_min} plane={cut_plane}: sigma+ MC {mcp:.6f} (+-{incp:.1e}) fawzi {sp:.6f} "
          f"ratio-1 {mcp / sp - 1:.2e}; sigma- MC {mcm:.6f} fawzi {sm:.6f} ratio-1 {mcm / sm - 1:.2e}")
    for mc, th in ((mcp, sp), (mcm, sm)):
        assert -2.0e-3 < mc / th - 1 < -0.6e-3


def test_oracle_plane_cut_dependence_matches_fawzi(oracle, valeurs):
    """The plane cut's effect alone, sigma(plane 0.3) / sigma(plane 0) on the
    same events, against the analytic ratio."""
    res = []
    for plane in (0.0, 0.3):
        cfg = oracle.config(valeurs, 4_000_000)
        cfg.cut_beam, cfg.cut_plane = 1.0, plane
        merged, _ = oracle.run(cfg, oracle.opts("xoshiro", threads=8))
        fin = oracle.finalize(cfg, merged)
        res.append((mc_sigma_pm(fin.spm2, fin.vars), fawzi_sigma(cfg)))
    for k in range(2):
        mc_ratio = res[1][0][k][0] / res[0][0][k][0]
        th_ratio = res[1][1][k] / res[0][1][k]
        print(f"sigma{'+-'[k]}: plane-cut ratio MC {mc_ratio:.6f} fawzi {th_ratio:.6f}")
        assert abs(mc_ratio / th_ratio - 1) < 2e-4


# ---- the GPU's per-event output under the same identities -------------------

@pytest.mark.gpu
def test_gpu_events_satisfy_rambo_and_berends_kleiss(valeurs):
    cfg = tp.Configuration.load(valeurs)
    n = 20000
    print(f"gpu invariants n={n}", flush=True)
    with tp.Engine(0) as eng:
        for kind in (tp.RNG_XOSHIRO, tp.RNG_RANF):
            st = tp.rng_state(kind)
            u, p, keep, m = eng.events(tp.params_from_config(cfg), kind, st, n)
            check_rambo(p, u, cfg.e_total)
            k = keep.astype(bool)
            full = np.zeros((n, 5, 4))
            full[:, 0] = [-cfg.e_total / 2, 0, 0, cfg.e_total / 2]
            full[:, 1] = [cfg.e_total / 2, 0, 0, cfg.e_total / 2]
            full[:, 2:] = p
            check_qed(full[k], m[k, 0], cfg.alpha)


@pytest.mark.gpu
def test_gpu_anomalous_sigma_against_fawzi(valeurs):
    """1e7 events on the GPU without the beam cut against fawzi(), as the
    oracle test above."""
    cfg = tp.Configuration.load(valeurs).with_events(10_000_000)
    cfg.raw.cut_beam = 1.0
    print("gpu fawzi run N=1e7", flush=True)
    with tp.Engine(0) as eng:
        res = tp.run_simulation(cfg, rng="xoshiro", engine=eng)
    (mcp, _), (mcm, _) = mc_sigma_pm(res.final.spm2, res.final.vars)
    sp, sm = fawzi_sigma(cfg)
    print(f"  sigma+ {mcp:.6f} fawzi {sp:.6f} ({mcp / sp - 1:.2e}); sigma- {mcm:.6f} fawzi {sm:.6f} ({mcm / sm - 1:.2e})")
    for mc, th in ((mcp, sp), (mcm, sm)):
        assert -2.0e-3 < mc / th - 1 < -0.6e-3
