"""The numbers res.data prints (output.rs:64-142), recomputed from a final
result in its own precision, in file order -- for comparing unrounded values
against the golden files' printed digits."""
from __future__ import annotations

import numpy as np


def printed_values(cfg, fin, f32: bool = False) -> list[float]:
    T = np.float32 if f32 else np.float64
    x = lambda v: T(v)
    one = T(1)
    vals = [cfg.num_events, fin.selected if hasattr(fin, "selected") else fin.selected_events,
            x(cfg.e_total), x(cfg.cut_beam), x(cfg.cut_pp), x(cfg.cut_plane), x(cfg.e_min),
            one / x(cfg.alpha), one / x(cfg.alpha_z), x(cfg.conv if hasattr(cfg, "conv") else cfg.gev2_to_picobarn),
            x(cfg.m_z0), x(cfg.g_z0), x(cfg.sin2w if hasattr(cfg, "sin2w") else cfg.sin2_weinberg),
            x(cfg.br_ep_em if hasattr(cfg, "br_ep_em") else cfg.branching_ep_em), x(cfg.beta_plus), x(cfg.beta_minus)]
    s, p = T(fin.sigma), T(fin.prec)
    vals += [s, s * p, p, T(fin.beta_min), T(fin.ss_p), T(fin.ss_p) * T(fin.inc_ss_p), T(fin.ss_m),
             T(fin.ss_m) * T(fin.inc_ss_m)]
    spm2 = np.array([[T(fin.spm2[sp][k]) for k in range(5)] for sp in range(2)], dtype=T)
    vars_ = np.array([[T(fin.vars[sp][k]) for k in range(5)] for sp in range(2)], dtype=T)
    for sp in range(2):
        for k in range(5):
            vals += [sp + 1, k + 1, spm2[sp, k], abs(spm2[sp, k]) * vars_[sp, k], vars_[sp, k]]
    for k in range(5):
        t1 = (T(0) + spm2[0, k]) + spm2[1, k]
        a, b = spm2[0, k] * vars_[0, k], spm2[1, k] * vars_[1, k]
        t2 = np.sqrt(T(0) + (a * a + b * b))
        with np.errstate(invalid="ignore", divide="ignore"):
            vals += [k + 1, t1 / T(4), t2 / T(4), t2 / abs(t1)]
    return [float(v) for v in vals]


def golden_numbers(text: str) -> list[str]:
    from numdiff import tokens
    return [t for k, t in tokens(text) if k == "num"]
