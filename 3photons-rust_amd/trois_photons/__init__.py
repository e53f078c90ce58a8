"""Python binding of the MI355X engine for the 3photons event loop.

Thin ctypes layer over ``build/libtp.so`` (C ABI in ``include/tp.h`` and
``include/tp_host.h``).  It mirrors the host side of the reference
(HadrienG2/3photons-rust): ``Configuration.load`` (src/config.rs:57-128),
``params_from_config`` (src/coupling.rs:23-33, src/resacc.rs:59-100),
``run_simulation`` (src/scheduling/mod.rs:31-59) and ``finalize``
(src/resacc.rs:142-223), with the event loop itself on the GPU.

There is no CPU fallback: if the HIP library is missing, importing this module
raises; if no GPU is usable, ``Engine()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "build", "libtp.so")

BATCH_SIZE = 10000
RNG_RANF, RNG_XOSHIRO, RNG_XOSHIRO128 = 0, 1, 2
SEED_SKIPAHEAD, SEED_JUMP = 0, 1
PREC_F64, PREC_MIXED, PREC_F32 = 0, 1, 2
TP_OK = 0

RNG_KINDS = {"ranf": RNG_RANF, "xoshiro": RNG_XOSHIRO}
SEEDINGS = {"skipahead": SEED_SKIPAHEAD, "jump": SEED_JUMP}


class TpParams(C.Structure):
    _fields_ = [("e_total", C.c_double), ("cut_beam", C.c_double), ("cut_pp", C.c_double),
                ("e_min", C.c_double), ("cut_plane", C.c_double), ("g_a", C.c_double),
                ("g_beta_p", C.c_double), ("g_beta_m", C.c_double), ("sigma_contribs", C.c_double * 5),
                ("ev_weight", C.c_double), ("sort_photons", C.c_int32), ("precision", C.c_int32),
                ("faster_evgen", C.c_int32), ("reserved", C.c_int32)]


class TpAcc(C.Structure):
    _fields_ = [("selected_events", C.c_uint64), ("spm2", C.c_double * 5), ("vars", C.c_double * 5),
                ("sigma", C.c_double), ("variance", C.c_double)]

    def as_array(self) -> np.ndarray:
        """[selected, spm2*5, vars*5, sigma, variance] as float64."""
        return np.array([self.selected_events, *self.spm2, *self.vars, self.sigma, self.variance],
                        dtype=np.float64)


ACC_DTYPE = np.dtype([("selected_events", np.uint64), ("spm2", np.float64, 5), ("vars", np.float64, 5),
                      ("sigma", np.float64), ("variance", np.float64)])
assert ACC_DTYPE.itemsize == C.sizeof(TpAcc) == 104


class RanfState(C.Structure):
    _fields_ = [("numbers", C.c_int32 * 56), ("index", C.c_uint32), ("seed", C.c_int32)]


class XoshiroState(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class Xoshiro128State(C.Structure):
    _fields_ = [("s", C.c_uint32 * 4)]


class TpStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("kernel_ms", C.c_double), ("events", C.c_uint64),
                ("batches", C.c_uint64), ("host_walk_chunks", C.c_uint64), ("chunks", C.c_uint64)]


class TpConfig(C.Structure):
    _fields_ = [("num_events", C.c_uint64)] + [
        (n, C.c_double) for n in ("e_total", "cut_beam", "cut_pp", "e_min", "cut_plane", "alpha", "alpha_z",
                                  "gev2_to_picobarn", "m_z0", "g_z0", "sin2_weinberg", "branching_ep_em",
                                  "beta_plus", "beta_minus")] + [
        (n, C.c_int32) for n in ("num_bins", "impr", "plot", "parsed")]


class TpFlowOpts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("batch", C.c_uint64), ("steps", C.c_int32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("adam_eps", C.c_float), ("grad_clip", C.c_float),
                ("eps_rel", C.c_double), ("cut_width", C.c_double), ("target", C.c_int32),
                ("cap_rel", C.c_double), ("seed", C.c_uint64)]


FLOW_NPARAM = 43968
FLOW_LAYER = 7328
FLOW_COND_MASKS = (0x03F, 0xFC0, 0x555, 0xAAA, 0x1C7, 0xE38)
UNIFORM_KEY_XOR = 0x5555555555555555


class TpFinal(C.Structure):
    _fields_ = [("selected_events", C.c_uint64), ("spm2", (C.c_double * 5) * 2), ("vars", (C.c_double * 5) * 2)] + [
        (n, C.c_double) for n in ("sigma", "prec", "variance", "beta_min", "ss_p", "inc_ss_p", "ss_m", "inc_ss_m")]


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libtp.so not built at {LIB_PATH}; run __graft_entry__.build() or make -C 3photons-rust_amd")
    # PyTorch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7) as the system one libtp.so links: whichever is loaded
    # first serves the whole process, and torch cannot use the system one.  So
    # when torch is installed it is loaded first, and libtp.so binds to its
    # runtime, whichever of the two modules the caller imports first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp, P = C.c_void_p, C.POINTER
    sig = {
        "tp_create": (C.c_int, [C.c_int, P(vp)]),
        "tp_destroy": (C.c_int, [vp]),
        "tp_last_error": (C.c_char_p, [vp]),
        "tp_rng_init": (C.c_int, [C.c_int, vp]),
        "tp_rng_seed": (C.c_int, [C.c_int, C.c_uint64, vp]),
        "tp_rng_skip": (C.c_int, [C.c_int, vp, C.c_uint64]),
        "tp_rng_jump": (C.c_int, [C.c_int, vp]),
        "tp_simulate_batch": (C.c_int, [vp, P(TpParams), C.c_int, vp, C.c_uint32, P(TpAcc)]),
        "tp_run": (C.c_int, [vp, P(TpParams), C.c_int, C.c_int, vp, C.c_uint64, vp, P(TpAcc), P(TpStats)]),
        "tp_run_device": (C.c_int, [vp, P(TpParams), C.c_int, C.c_int, vp, C.c_uint64, vp, vp, P(TpStats)]),
        "tp_run_device_noalloc": (C.c_int, [vp, P(TpParams), C.c_int, C.c_int, vp, C.c_uint64, vp, vp,
                                            P(TpStats)]),
        "tp_events": (C.c_int, [vp, P(TpParams), C.c_int, vp, C.c_uint32, vp, vp, vp, vp]),
        "tp_adjoint": (C.c_int, [vp, P(TpParams), C.c_int, vp, vp, C.c_uint64, C.c_int, C.c_int, vp, vp, vp]),
        "tp_sync": (C.c_int, [vp, P(TpStats)]),
        "tp_flow_default_opts": (C.c_int, [P(TpFlowOpts)]),
        "tp_flow_create": (C.c_int, [vp, C.c_uint64, P(vp)]),
        "tp_flow_destroy": (C.c_int, [vp]),
        "tp_flow_get_params": (C.c_int, [vp, vp]),
        "tp_flow_set_params": (C.c_int, [vp, vp]),
        "tp_flow_train": (C.c_int, [vp, P(TpParams), P(TpFlowOpts), vp, P(TpStats)]),
        "tp_flow_sample": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, C.c_int, vp, vp, vp]),
        "tp_flow_density": (C.c_int, [vp, C.c_uint64, vp, C.c_int, vp]),
        "tp_flow_grad": (C.c_int, [vp, C.c_uint64, vp, vp, C.c_int, vp, vp]),
        "tp_flow_integrand": (C.c_int, [vp, P(TpParams), C.c_uint64, vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_float, C.c_int, vp, vp, vp]),
        "tp_flow_run_device": (C.c_int, [vp, P(TpParams), C.c_double, C.c_uint64, C.c_uint64, vp, vp,
                                         P(TpStats)]),
        "tp_flow_run": (C.c_int, [vp, P(TpParams), C.c_double, C.c_uint64, C.c_uint64, vp, P(TpAcc), P(TpStats)]),
        "tp_config_load": (C.c_int, [C.c_char_p, P(TpConfig), C.c_char_p, C.c_int]),
        "tp_config_format": (C.c_int, [P(TpConfig), C.c_char_p, C.c_int]),
        "tp_params_from_config": (C.c_int, [P(TpConfig), P(TpParams)]),
        "tp_acc_merge": (None, [P(TpAcc), P(TpAcc)]),
        "tp_finalize": (C.c_int, [P(TpConfig), P(TpAcc), P(TpFinal)]),
        "tp_params_from_config_f32": (C.c_int, [P(TpConfig), P(TpParams)]),
        "tp_config_format_f32": (C.c_int, [P(TpConfig), C.c_char_p, C.c_int]),
        "tp_format_res_data_f32": (C.c_int, [P(TpConfig), P(TpFinal), C.c_char_p, C.c_int]),
        "tp_format_report_f32": (C.c_int, [P(TpConfig), P(TpFinal), C.c_char_p, C.c_int]),
        "tp_dump_results_f32": (C.c_int, [P(TpConfig), P(TpFinal), C.c_double, C.c_char_p]),
        "tp_acc_merge_f32": (None, [P(TpAcc), P(TpAcc)]),
        "tp_finalize_f32": (C.c_int, [P(TpConfig), P(TpAcc), P(TpFinal)]),
        "tp_format_res_data": (C.c_int, [P(TpConfig), P(TpFinal), C.c_char_p, C.c_int]),
        "tp_format_report": (C.c_int, [P(TpConfig), P(TpFinal), C.c_char_p, C.c_int]),
        "tp_dump_results": (C.c_int, [P(TpConfig), P(TpFinal), C.c_double, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class TpError(RuntimeError):
    pass


# ---- host side --------------------------------------------------------------

class Configuration:
    """The 18 items of `valeurs` (config.rs:8-53)."""

    def __init__(self, raw: TpConfig):
        self.raw = raw

    @classmethod
    def load(cls, path: str) -> "Configuration":
        raw = TpConfig()
        err = C.create_string_buffer(1024)
        rc = lib.tp_config_load(path.encode(), C.byref(raw), err, len(err))
        if rc != TP_OK:
            raise TpError(err.value.decode())
        return cls(raw)

    def __getattr__(self, name):
        return getattr(self.raw, name)

    def with_events(self, n: int) -> "Configuration":
        raw = TpConfig.from_buffer_copy(self.raw)
        raw.num_events = n
        return Configuration(raw)

    def display(self, f32: bool = False) -> str:
        return _text(lib.tp_config_format_f32 if f32 else lib.tp_config_format, C.byref(self.raw))


def _text(fn, *args) -> str:
    n = fn(*args, None, 0)
    buf = C.create_string_buffer(n)
    fn(*args, buf, n)
    return buf.value.decode()


def params_from_config(cfg: Configuration, sort: bool = True, precision: int = PREC_F64,
                       faster_evgen: bool = False) -> TpParams:
    """Run constants; precision PREC_F32 computes them in f32 like the
    reference's f32 feature (tp_params_from_config_f32)."""
    p = TpParams()
    fn = lib.tp_params_from_config_f32 if precision == PREC_F32 else lib.tp_params_from_config
    rc = fn(C.byref(cfg.raw), C.byref(p))
    if rc != TP_OK:
        raise TpError(f"tp_params_from_config: {rc}")
    p.sort_photons = 1 if sort else 0
    p.precision = precision
    p.faster_evgen = 1 if faster_evgen else 0
    return p


def rng_state(kind: int, seed: int | None = None):
    st = {RNG_RANF: RanfState, RNG_XOSHIRO: XoshiroState, RNG_XOSHIRO128: Xoshiro128State}[kind]()
    rc = lib.tp_rng_init(kind, C.byref(st)) if seed is None else lib.tp_rng_seed(kind, seed, C.byref(st))
    if rc != TP_OK:
        raise TpError(f"rng init: {rc}")
    return st


def rng_skip(kind: int, st, n: int) -> None:
    if lib.tp_rng_skip(kind, C.byref(st), n) != TP_OK:
        raise TpError("tp_rng_skip")


def rng_jump(kind: int, st) -> None:
    if lib.tp_rng_jump(kind, C.byref(st)) != TP_OK:
        raise TpError("tp_rng_jump")


def merge(accs, f32: bool = False) -> TpAcc:
    """Left fold in batch order (resacc.rs:133-139), in f32 for f32 runs."""
    fn = lib.tp_acc_merge_f32 if f32 else lib.tp_acc_merge
    out = TpAcc.from_buffer_copy(accs[0])
    for a in accs[1:]:
        if not isinstance(a, TpAcc):  # element of an ACC_DTYPE array
            a = TpAcc.from_buffer_copy(a)
        fn(C.byref(out), C.byref(a))
    return out


def finalize(cfg: Configuration, acc: TpAcc, f32: bool = False) -> TpFinal:
    f = TpFinal()
    fn = lib.tp_finalize_f32 if f32 else lib.tp_finalize
    if fn(C.byref(cfg.raw), C.byref(acc), C.byref(f)) != TP_OK:
        raise TpError("tp_finalize")
    return f


def format_res_data(cfg: Configuration, res: TpFinal, f32: bool = False) -> str:
    return _text(lib.tp_format_res_data_f32 if f32 else lib.tp_format_res_data, C.byref(cfg.raw), C.byref(res))


def format_report(cfg: Configuration, res: TpFinal, f32: bool = False) -> str:
    return _text(lib.tp_format_report_f32 if f32 else lib.tp_format_report, C.byref(cfg.raw), C.byref(res))


# ---- engine -------------------------------------------------------------------

class Engine:
    """One tp context: one GPU, one stream (tp_create / tp_destroy)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.tp_create(device, C.byref(h))
        if rc != TP_OK:
            raise TpError(f"tp_create(device={device}) failed with status {rc}: no usable GPU")
        self.h = h
        self._flows = weakref.WeakSet()  # flows use the context: closed first

    def close(self):
        if self.h:
            for f in list(self._flows):
                f.close()
            lib.tp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != TP_OK:
            raise TpError(f"{what}: status {rc}: {lib.tp_last_error(self.h).decode()}")

    def simulate_batch(self, params: TpParams, kind: int, st, n: int) -> TpAcc:
        """One call of the reference's per-batch closure; advances st."""
        out = TpAcc()
        self._check(lib.tp_simulate_batch(self.h, C.byref(params), kind, C.byref(st), n, C.byref(out)),
                    "tp_simulate_batch")
        return out

    def run(self, params: TpParams, kind: int, seeding: int, st, num_events: int):
        """Whole run: (per-batch accumulators as a structured array, merged TpAcc, TpStats)."""
        nb = (num_events + BATCH_SIZE - 1) // BATCH_SIZE
        per = np.zeros(nb, dtype=ACC_DTYPE)
        merged, stats = TpAcc(), TpStats()
        self._check(lib.tp_run(self.h, C.byref(params), kind, seeding, C.byref(st), num_events,
                               per.ctypes.data_as(C.c_void_p), C.byref(merged), C.byref(stats)), "tp_run")
        return per, merged, stats

    def run_device(self, params: TpParams, kind: int, seeding: int, st, num_events: int,
                   per_batch_ptr: int, merged_ptr: int | None) -> None:
        self._check(lib.tp_run_device(self.h, C.byref(params), kind, seeding, C.byref(st), num_events,
                                      C.c_void_p(per_batch_ptr), C.c_void_p(merged_ptr or 0), None),
                    "tp_run_device")

    def run_device_noalloc(self, params: TpParams, kind: int, seeding: int, st, num_events: int,
                           per_batch_ptr: int, merged_ptr: int | None) -> None:
        """tp_run_device_noalloc: the same run, tables from the static device array."""
        self._check(lib.tp_run_device_noalloc(self.h, C.byref(params), kind, seeding, C.byref(st), num_events,
                                              C.c_void_p(per_batch_ptr), C.c_void_p(merged_ptr or 0), None),
                    "tp_run_device_noalloc")

    def sync(self) -> TpStats:
        s = TpStats()
        self._check(lib.tp_sync(self.h, C.byref(s)), "tp_sync")
        return s

    def adjoint(self, params: TpParams, n: int, precision: int = PREC_F64, kind: int = RNG_XOSHIRO, st=None,
                u: np.ndarray | None = None):
        """F[n,5], dF/du[n,5,12], dF/dp[n,5,12] (host arrays) from tp_adjoint."""
        dt = np.float32 if precision == PREC_F32 else np.float64
        F = np.zeros((n, 5), dt); gu = np.zeros((n, 5, 12), dt); gp = np.zeros((n, 5, 12), dt)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        uin = None
        if u is not None:
            uin = np.ascontiguousarray(u, dtype=np.float64).reshape(n, 12)
        self._check(lib.tp_adjoint(self.h, C.byref(params), kind, C.byref(st) if st is not None else None,
                                   ptr(uin) if uin is not None else None, n, precision, 0, ptr(F), ptr(gu), ptr(gp)),
                    "tp_adjoint")
        return F, gu, gp

    def adjoint_device(self, params: TpParams, n: int, precision: int, st, F_ptr: int, gu_ptr: int, gp_ptr: int):
        self._check(lib.tp_adjoint(self.h, C.byref(params), RNG_XOSHIRO, C.byref(st), None, n, precision, 1,
                                   C.c_void_p(F_ptr), C.c_void_p(gu_ptr), C.c_void_p(gp_ptr)), "tp_adjoint")

    def last_stats(self) -> TpStats:
        return self.sync()

    def events(self, params: TpParams, kind: int, st, n: int):
        """Per-event diagnostics: (u[n,12], p[n,3,4], keep[n], m[n,5])."""
        u = np.zeros((n, 12)); p = np.zeros((n, 3, 4)); k = np.zeros(n, dtype=np.int32); m = np.zeros((n, 5))
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(lib.tp_events(self.h, C.byref(params), kind, C.byref(st), n, ptr(u), ptr(p), ptr(k), ptr(m)),
                    "tp_events")
        return u, p, k, m


def flow_default_opts(**kw) -> TpFlowOpts:
    o = TpFlowOpts()
    lib.tp_flow_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Flow:
    """RealNVP proposal on the 12 RAMBO uniforms, bound to an Engine
    (tp_flow_* entry points).  Host arrays in and out unless a method says
    otherwise."""

    def __init__(self, engine: Engine, init_seed: int = 0):
        h = C.c_void_p()
        self.engine = engine
        engine._check(lib.tp_flow_create(engine.h, init_seed, C.byref(h)), "tp_flow_create")
        self.h = h
        engine._flows.add(self)

    def close(self):
        if self.h:
            lib.tp_flow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        self.engine._check(rc, what)

    def get_params(self) -> np.ndarray:
        w = np.zeros(FLOW_NPARAM, np.float32)
        self._check(lib.tp_flow_get_params(self.h, _ptr(w)), "tp_flow_get_params")
        return w

    def set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == FLOW_NPARAM
        self._check(lib.tp_flow_set_params(self.h, _ptr(w)), "tp_flow_set_params")

    def train(self, params: TpParams, opts: TpFlowOpts):
        """(loss per step, TpStats)."""
        loss = np.zeros(max(1, opts.steps), np.float32)
        st = TpStats()
        self._check(lib.tp_flow_train(self.h, C.byref(params), C.byref(opts), _ptr(loss), C.byref(st)),
                    "tp_flow_train")
        return loss[:opts.steps], st

    def sample(self, n: int, key: int = 0, sample0: int = 0, v: np.ndarray | None = None):
        """(u[n,12] f64, z[n,12] f32, log q[n] f32)."""
        u = np.zeros((n, 12)); z = np.zeros((n, 12), np.float32); lq = np.zeros(n, np.float32)
        vin = None if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(n, 16)
        self._check(lib.tp_flow_sample(self.h, key, sample0, n, _ptr(vin), 0, _ptr(u), _ptr(z), _ptr(lq)),
                    "tp_flow_sample")
        return u, z, lq

    def density(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 12)
        lq = np.zeros(len(u), np.float32)
        self._check(lib.tp_flow_density(self.h, len(u), _ptr(u), 0, _ptr(lq)), "tp_flow_density")
        return lq

    def grad(self, v: np.ndarray, gz: np.ndarray):
        """(d/dparams, d/dz0 per sample) of (1/n) sum_i [gz_i . z_i - sum s]."""
        n = len(v)
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(n, 16)
        gz = np.ascontiguousarray(gz, dtype=np.float32).reshape(n, 16)
        g = np.zeros(FLOW_NPARAM, np.float32); gv = np.zeros((n, 16), np.float32)
        self._check(lib.tp_flow_grad(self.h, n, _ptr(v), _ptr(gz), 0, _ptr(g), _ptr(gv)), "tp_flow_grad")
        return g, gv

    def integrand(self, params: TpParams, u: np.ndarray, eps: float, tau: float = 0.0, target: int = 0,
                  clip: float = 1e30, fcap: float = float("inf")):
        """(f[n] f64, gz[n,16] f32, log g[n] f32) of the training target (tp_flow_integrand)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 12)
        n = len(u)
        f = np.zeros(n); gz = np.zeros((n, 16), np.float32); lf = np.zeros(n, np.float32)
        self._check(lib.tp_flow_integrand(self.h, C.byref(params), n, _ptr(u), eps, tau, fcap, target, clip, 0,
                                          _ptr(f), _ptr(gz), _ptr(lf)), "tp_flow_integrand")
        return f, gz, lf

    def run(self, params: TpParams, alpha: float, key: int, num_events: int):
        """Weighted run: (per-batch accumulators, merged TpAcc, TpStats), host side."""
        nb = (num_events + BATCH_SIZE - 1) // BATCH_SIZE
        per = np.zeros(nb, dtype=ACC_DTYPE)
        merged, stats = TpAcc(), TpStats()
        self._check(lib.tp_flow_run(self.h, C.byref(params), alpha, key, num_events, per.ctypes.data_as(C.c_void_p),
                                    C.byref(merged), C.byref(stats)), "tp_flow_run")
        return per, merged, stats

    def run_device(self, params: TpParams, alpha: float, key: int, num_events: int, per_batch_ptr: int,
                   merged_ptr: int | None) -> None:
        self._check(lib.tp_flow_run_device(self.h, C.byref(params), alpha, key, num_events,
                                           C.c_void_p(per_batch_ptr), C.c_void_p(merged_ptr or 0), None),
                    "tp_flow_run_device")


@dataclass
class RunResult:
    per_batch: np.ndarray
    merged: TpAcc
    final: TpFinal
    stats: TpStats


def run_simulation(cfg: Configuration, rng: str = "ranf", seeding: str = "skipahead", sort: bool = True,
                   precision: int = PREC_F64, engine: Engine | None = None,
                   faster_evgen: bool = False) -> RunResult:
    """scheduling::run_simulation (scheduling/mod.rs:31-59) on the GPU."""
    kind = RNG_KINDS[rng]
    if kind == RNG_XOSHIRO and precision == PREC_F32:
        kind = RNG_XOSHIRO128  # the f32 feature's generator (standard.rs:8-11)
    params = params_from_config(cfg, sort=sort, precision=precision, faster_evgen=faster_evgen)
    st = rng_state(kind)
    own = engine is None
    eng = engine or Engine()
    try:
        per, merged, stats = eng.run(params, kind, SEEDINGS[seeding], st, cfg.num_events)
    finally:
        if own:
            eng.close()
    return RunResult(per, merged, finalize(cfg, merged, f32=precision == PREC_F32), stats)
