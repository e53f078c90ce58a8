"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests, the smoke check and bench.py's CPU-baseline
leg use it, and only as the checker / the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle", "build")


def _scalar(f32: bool):
    return C.c_float if f32 else C.c_double


class OrConfig(C.Structure):
    _fields_ = [("num_events", C.c_uint64)] + [
        (n, C.c_double) for n in ("e_total", "cut_beam", "cut_pp", "e_min", "cut_plane", "alpha", "alpha_z", "conv",
                                  "m_z0", "g_z0", "sin2w", "br_ep_em", "beta_plus", "beta_minus")] + [
        (n, C.c_int32) for n in ("num_bins", "impr", "plot", "pad")]


class OrOpts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("rng_kind", "faster_threading", "sort", "threads", "faster_evgen", "pad")]


class OrRng(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ranf_seed", C.c_int32), ("numbers", C.c_int32 * 56), ("index", C.c_uint32),
                ("pad", C.c_uint32), ("xs", C.c_uint64 * 4), ("xs32", C.c_uint32 * 4)]


def _acc_type(f32):
    S = _scalar(f32)

    class OrAcc(C.Structure):
        _fields_ = [("selected", C.c_uint64), ("spm2", S * 5), ("vars", S * 5), ("sigma", S), ("variance", S)]

        def as_array(self):
            return np.array([self.selected, *self.spm2, *self.vars, self.sigma, self.variance], dtype=np.float64)
    return OrAcc


def _final_type(f32):
    S = _scalar(f32)

    class OrFinal(C.Structure):
        _fields_ = [("selected", C.c_uint64), ("spm2", (S * 5) * 2), ("vars", (S * 5) * 2)] + [
            (n, S) for n in ("sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m")]
    return OrFinal


class Oracle:
    def __init__(self, f32: bool = False, path: str | None = None):
        self.f32 = f32
        path = path or os.path.join(ORACLE_DIR, "liboracle_f32.so" if f32 else "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle not built: {path} (run make -C oracle)")
        self.lib = C.CDLL(path)
        self.S = _scalar(f32)
        self.Acc = _acc_type(f32)
        self.Final = _final_type(f32)
        sfx = "_f32" if f32 else ""
        self._fn = lambda name: getattr(self.lib, name + sfx)
        self._fn("or_rng_random").restype = self.S
        self._fn("or_rng_next_raw").restype = C.c_uint64

    # -- configuration -------------------------------------------------------
    def config(self, path: str, num_events: int | None = None) -> OrConfig:
        cfg = OrConfig()
        err = C.create_string_buffer(256)
        rc = self._fn("or_config_load")(path.encode(), C.byref(cfg), err, 256)
        if rc != 0:
            raise ValueError(err.value.decode())
        if num_events is not None:
            cfg.num_events = num_events
        return cfg

    @staticmethod
    def opts(rng="ranf", faster_threading=False, sort=True, threads=1, faster_evgen=False) -> OrOpts:
        return OrOpts({"ranf": 0, "xoshiro": 1}[rng], int(faster_threading), int(sort), threads, int(faster_evgen), 0)

    # -- streams ---------------------------------------------------------------
    def rng(self, kind: str = "ranf", seed: int | None = None) -> OrRng:
        g = OrRng()
        k = {"ranf": 0, "xoshiro": 1}[kind]
        if seed is None:
            self._fn("or_rng_new")(k, C.byref(g))
        else:
            self._fn("or_rng_seeded")(k, C.c_uint64(seed), C.byref(g))
        return g

    def uniforms(self, g: OrRng, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.float32 if self.f32 else np.float64)
        self._fn("or_rng_fill")(C.byref(g), C.c_uint64(n), out.ctypes.data_as(C.c_void_p))
        return out

    def raw(self, g: OrRng, n: int) -> np.ndarray:
        f = self._fn("or_rng_next_raw")
        return np.array([f(C.byref(g)) for _ in range(n)], dtype=np.uint64)

    def skip(self, g: OrRng, n: int):
        self._fn("or_rng_skip")(C.byref(g), C.c_uint64(n))

    def jump(self, g: OrRng):
        self._fn("or_rng_jump")(C.byref(g))

    # -- events, batches, runs -------------------------------------------------
    def events(self, cfg: OrConfig, opts: OrOpts, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float32 if self.f32 else np.float64).reshape(-1, 12)
        n = u.shape[0]
        dt = u.dtype
        p = np.zeros((n, 5, 4), dt); keep = np.zeros(n, np.int32); m2 = np.zeros((n, 5, 8), dt)
        m = np.zeros((n, 5), dt); w = np.zeros(n, dt)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        self._fn("or_events_from_uniforms")(C.byref(cfg), C.byref(opts), C.c_uint64(n), ptr(u), ptr(p), ptr(keep),
                                            ptr(m2), ptr(m), ptr(w))
        return p, keep.astype(bool), m2, m, w

    def spinors(self, p: np.ndarray) -> np.ndarray:
        """s_ij of given momenta p[n][5][4] as complex [n][5][5]."""
        dt = np.float32 if self.f32 else np.float64
        p = np.ascontiguousarray(p, dtype=dt).reshape(-1, 5, 4)
        s = np.zeros((p.shape[0], 5, 5, 2), dt)
        self._fn("or_spinors_at")(C.c_uint64(p.shape[0]), p.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p))
        return s[..., 0] + 1j * s[..., 1]

    def simulate_batch(self, cfg, opts, g: OrRng, n: int):
        acc = self.Acc()
        self._fn("or_simulate_batch")(C.byref(cfg), C.byref(opts), C.byref(g), C.c_uint64(n), C.byref(acc))
        return acc

    def run(self, cfg, opts, per_batch: bool = False):
        nb = (cfg.num_events + 9999) // 10000
        merged = self.Acc()
        per = (self.Acc * nb)() if per_batch else None
        rc = self._fn("or_run")(C.byref(cfg), C.byref(opts), C.byref(merged), per if per_batch else None)
        if rc != 0:
            raise RuntimeError(f"or_run: {rc}")
        return merged, per

    def finalize(self, cfg, acc):
        f = self.Final()
        self._fn("or_finalize")(C.byref(cfg), C.byref(acc), C.byref(f))
        return f


class OracleGrad:
    """Forward-mode (dual number) f64 gradients of the oracle's event
    integrand (oracle/or_grad.cpp)."""

    def __init__(self, oracle: Oracle, cfg: OrConfig, sort: bool = True):
        path = os.path.join(ORACLE_DIR, "liborgrad.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle gradient not built: {path}")
        self.lib = C.CDLL(path)

        class K(C.Structure):
            _fields_ = [(n, C.c_double) for n in ("ev_weight", "g_a", "g_bp", "g_bm", "fact_com", "norm_weight",
                                                   "propagator", "delta_z0")] + [("sc", C.c_double * 5)]
        k = K()
        oracle.lib.or_consts(C.byref(cfg), C.byref(k))
        self.k = np.array([cfg.e_total, cfg.cut_beam, cfg.cut_pp, cfg.e_min, cfg.cut_plane, k.g_a, k.g_bp, k.g_bm,
                           k.ev_weight], dtype=np.float64)
        self.sort = int(sort)

    def grad(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 12)
        n = u.shape[0]
        F = np.zeros((n, 5)); gu = np.zeros((n, 5, 12)); gp = np.zeros((n, 5, 12))
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        self.lib.or_grad(ptr(self.k), self.sort, C.c_uint64(n), ptr(u), ptr(F), ptr(gu), ptr(gp))
        return F, gu, gp

    def m2_at(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 5, 4)
        n = p.shape[0]
        F = np.zeros((n, 5))
        self.lib.or_m2_at(self.k.ctypes.data_as(C.c_void_p), C.c_uint64(n), p.ctypes.data_as(C.c_void_p),
                          F.ctypes.data_as(C.c_void_p))
        return F
