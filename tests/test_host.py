"""Product host side (libtp.so, tp_host.h) on the CPU: res.data and stdout
formats against the shipped files, configuration parsing and its errors,
finalize against the oracle."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import trois_photons as tp
from conftest import PKG, golden_path
from numdiff import compare_texts


def to_tp_acc(or_acc) -> tp.TpAcc:
    a = tp.TpAcc()
    a.selected_events = or_acc.selected
    for i in range(5):
        a.spm2[i] = or_acc.spm2[i]
        a.vars[i] = or_acc.vars[i]
    a.sigma, a.variance = or_acc.sigma, or_acc.variance
    return a


GOLDEN_RUNS = [("features_", "ranf", False, True, False), ("features_standard-random", "xoshiro", False, True, False),
               ("features_multi-threading,faster-threading", "ranf", True, True, False),
               ("features_faster-evgen", "ranf", False, True, True)]


@pytest.mark.parametrize("suffix,rng,ft,sort,fe", GOLDEN_RUNS)
def test_res_data_bytes_match_golden(oracle_run, valeurs, suffix, rng, ft, sort, fe):
    """Oracle accumulator -> product finalize + res.data writer == shipped bytes."""
    _, merged, _, _ = oracle_run(rng, ft, sort, fe)
    cfg = tp.Configuration.load(valeurs)
    fin = tp.finalize(cfg, to_tp_acc(merged))
    text = tp.format_res_data(cfg, fin)
    golden = open(golden_path("res.data-" + suffix)).read()
    assert compare_texts(golden, text) == []
    assert text == golden


@pytest.mark.parametrize("suffix,rng,ft,sort,fe", GOLDEN_RUNS)
def test_stdout_matches_golden(oracle_run, valeurs, suffix, rng, ft, sort, fe):
    _, merged, _, _ = oracle_run(rng, ft, sort, fe)
    cfg = tp.Configuration.load(valeurs)
    fin = tp.finalize(cfg, to_tp_acc(merged))
    text = cfg.display() + "IBegin\n" + tp.format_report(cfg, fin)
    golden = open(golden_path("stdout.log-" + suffix)).read()
    assert compare_texts(golden, text) == []
    assert text == golden


def test_finalize_matches_oracle(oracle_run, valeurs):
    ocfg, merged, _, ofin = oracle_run("ranf")
    cfg = tp.Configuration.load(valeurs)
    fin = tp.finalize(cfg, to_tp_acc(merged))
    for f in ("sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m"):
        assert getattr(fin, f) == getattr(ofin, f), f
    assert np.array_equal(np.array(fin.spm2), np.array(ofin.spm2))


def test_params_match_oracle_constants(oracle, valeurs):
    """Couplings and sigma_contribs: product host == oracle, bit for bit."""
    cfg = tp.Configuration.load(valeurs)
    p = tp.params_from_config(cfg)
    from oracle_ffi import OrConfig  # noqa: F401
    import ctypes
    class K(ctypes.Structure):
        _fields_ = [(n, ctypes.c_double) for n in ("ev_weight", "g_a", "g_bp", "g_bm", "fact_com", "norm_weight",
                                                    "propagator", "delta_z0")] + [("sc", ctypes.c_double * 5)]
    k = K()
    oracle.lib.or_consts(ctypes.byref(oracle.config(valeurs)), ctypes.byref(k))
    assert (p.g_a, p.g_beta_p, p.g_beta_m) == (k.g_a, k.g_bp, k.g_bm)
    assert list(p.sigma_contribs) == list(k.sc)
    # closed form of the RAMBO weight (evgen.rs:40-77): s pi^2 / 8
    s = cfg.e_total ** 2
    assert abs(k.ev_weight / (s * np.pi ** 2 / 8) - 1) < 4e-15


def write(tmp_path, text, name="valeurs"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_config_display_matches_golden_head(valeurs):
    cfg = tp.Configuration.load(valeurs)
    golden = open(golden_path("stdout.log-features_")).read()
    assert golden.startswith(cfg.display())


@pytest.mark.parametrize("mutate,msg", [
    (lambda l: l[:5], "missing configuration of beam_photon_plane_cut"),
    (lambda l: ["abc"] + l[1:], "could not parse configuration of num_events"),
    (lambda l: [l[0], "91.x"] + l[2:], "could not parse configuration of e_total"),
    (lambda l: ["0"] + l[1:], "Please simulate at least one event"),
    (lambda l: l[:17] + [".true."] + l[18:], "Plotting is not supported"),
    (lambda l: l[:16] + [".TRUE."] + l[17:], "Individual result printing is not supported"),
    (lambda l: l[:16] + ["maybe"] + l[17:], "could not parse configuration of impr"),
])
def test_config_errors(tmp_path, valeurs, mutate, msg):
    lines = [l for l in open(valeurs).read().splitlines() if l.strip()][:18]
    path = write(tmp_path, "\n".join(mutate(lines)) + "\n")
    with pytest.raises(tp.TpError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        tp.Configuration.load(path)


def test_config_missing_file(tmp_path):
    with pytest.raises(tp.TpError, match="cannot open"):
        tp.Configuration.load(str(tmp_path / "nope"))


def test_config_blank_lines_and_extra_tokens(tmp_path, valeurs):
    lines = [l for l in open(valeurs).read().splitlines() if l.strip()][:18]
    text = "\n\n   \n".join(lines) + "\n"
    cfg = tp.Configuration.load(write(tmp_path, text))
    ref = tp.Configuration.load(valeurs)
    assert cfg.display() == ref.display()


def test_cli_config_errors(tmp_path, valeurs):
    exe = os.path.join(PKG, "build", "trois_photons")
    r = subprocess.run([exe, "--config", str(tmp_path / "missing")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    lines = [l for l in open(valeurs).read().splitlines() if l.strip()][:18]
    bad = write(tmp_path, "\n".join(["0"] + lines[1:]) + "\n")
    r = subprocess.run([exe, "--config", bad], capture_output=True, text=True)
    assert r.returncode == 1 and "Please simulate at least one event" in r.stderr
    assert r.stdout.startswith("ITOT           : 0\n")  # echoed before the check, like config.rs:108
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 2


# ---- all-f32 mode host side (the reference's f32 feature) ---------------------

def test_f32_params_match_oracle_constants(oracle32, valeurs):
    """tp_params_from_config_f32: the couplings and sigma_contribs computed in
    f32 arithmetic, bit for bit the f32 oracle's."""
    import ctypes
    cfg = tp.Configuration.load(valeurs)
    p = tp.params_from_config(cfg, precision=tp.PREC_F32)
    assert p.precision == tp.PREC_F32

    class K(ctypes.Structure):
        _fields_ = [(n, ctypes.c_float) for n in ("ev_weight", "g_a", "g_bp", "g_bm", "fact_com", "norm_weight",
                                                   "propagator", "delta_z0")] + [("sc", ctypes.c_float * 5)]
    k = K()
    oracle32.lib.or_consts_f32(ctypes.byref(oracle32.config(valeurs)), ctypes.byref(k))
    assert (p.g_a, p.g_beta_p, p.g_beta_m) == (k.g_a, k.g_bp, k.g_bm)
    assert list(p.sigma_contribs) == list(k.sc)
    assert p.e_total == float(np.float32(cfg.e_total))


def test_f32_merge_and_finalize_match_oracle(oracle32, valeurs):
    """f32 fold of per-batch accumulators and f32 finalize against the f32
    oracle's own, bit for bit, on a 2e5-event run."""
    ocfg = oracle32.config(valeurs, 200_000)
    merged, per = oracle32.run(ocfg, oracle32.opts("ranf", threads=4), per_batch=True)
    cfg = tp.Configuration.load(valeurs).with_events(200_000)
    m = tp.merge([to_tp_acc(a) for a in per], f32=True)
    assert m.as_array().tolist() == merged.as_array().tolist()
    fin = tp.finalize(cfg, m, f32=True)
    ofin = oracle32.finalize(ocfg, merged)
    for f in ("sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m"):
        assert getattr(fin, f) == getattr(ofin, f), f
    assert np.array_equal(np.array(fin.spm2), np.array(ofin.spm2))
    assert np.array_equal(np.array(fin.vars), np.array(ofin.vars))


@pytest.fixture(scope="session")
def oracle32_run(oracle32, valeurs):
    import os
    cache = {}

    def run(rng):
        if rng not in cache:
            ocfg = oracle32.config(valeurs)
            merged, _ = oracle32.run(ocfg, oracle32.opts(rng, threads=os.cpu_count() or 8))
            cache[rng] = merged
        return cache[rng]
    return run


@pytest.mark.parametrize("suffix,rng", [("features_f32", "ranf"), ("features_standard-random,f32", "xoshiro")])
def test_f32_outputs_match_golden_bytes(oracle32_run, valeurs, suffix, rng):
    """f32 oracle accumulator -> product f32 finalize and f32 writers: the
    reference's f32 res.data and stdout, byte for byte."""
    merged = oracle32_run(rng)
    cfg = tp.Configuration.load(valeurs)
    fin = tp.finalize(cfg, to_tp_acc(merged), f32=True)
    text = tp.format_res_data(cfg, fin, f32=True)
    golden = open(golden_path("res.data-" + suffix)).read()
    assert compare_texts(golden, text) == []
    assert text == golden
    out = cfg.display(f32=True) + "IBegin\n" + tp.format_report(cfg, fin, f32=True)
    golden = open(golden_path("stdout.log-" + suffix)).read()
    # The report prints some sums to 6 digits, where the f32 sums of 1e7
    # events carry the last ulps of the platform's sinf/cosf/logf: one unit in
    # the last printed digit is allowed there (features_f32 line 37 prints
    # Sig_Num 4.18073 where the oracle's f32 sums give 4.18074).
    assert compare_texts(golden, out, digits_slack=1.5) == []
