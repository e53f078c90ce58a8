"""Argument errors at the C-ABI boundary, on a live context: each entry point
returns TP_ERR_ARG (or TP_ERR_UNSUPPORTED) with a message, leaves the caller's
outputs and generator state untouched, and leaves the context usable -- a
valid call after the errors gives bit for bit the result it gave before them.
The reference's own checks are the config error "Please simulate at least
one event" (config.rs:131-133) and the closure's empty accumulator for a
batch of 0 events (evgen.rs:221-249 runs no event)."""
import ctypes as C

import numpy as np
import pytest

import trois_photons as tp

pytestmark = pytest.mark.gpu

lib = tp.lib
ERR_ARG, ERR_UNSUPPORTED = -1, -4


@pytest.fixture(scope="module")
def engine():
    print("creating engine on cuda:0", flush=True)
    with tp.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def params(valeurs):
    return tp.params_from_config(tp.Configuration.load(valeurs))


def last_error(engine) -> str:
    return lib.tp_last_error(engine.h).decode()


def expect(engine, rc, code, text):
    assert rc == code, (rc, last_error(engine))
    assert text in last_error(engine), last_error(engine)


def test_run_argument_errors_leave_the_context_usable(engine, params):
    n = 123_457
    print(f"reference run N={n}", flush=True)
    _, ref, _ = engine.run(params, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), n)
    per = np.zeros(13, dtype=tp.ACC_DTYPE)
    pp = per.ctypes.data_as(C.c_void_p)
    merged = tp.TpAcc()
    st = tp.rng_state(tp.RNG_XOSHIRO)
    words = bytes(st)
    print("tp_run argument errors", flush=True)
    expect(engine, lib.tp_run(engine.h, C.byref(params), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), 0, pp,
                              C.byref(merged), None), ERR_ARG, "at least one event")
    expect(engine, lib.tp_run(engine.h, C.byref(params), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), n, None,
                              C.byref(merged), None), ERR_ARG, "null per-batch output")
    expect(engine, lib.tp_run(engine.h, None, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), n, pp, C.byref(merged),
                              None), ERR_ARG, "null params")
    expect(engine, lib.tp_run(engine.h, C.byref(params), 7, tp.SEED_SKIPAHEAD, C.byref(st), n, pp, C.byref(merged),
                              None), ERR_ARG, "unknown rng kind")
    expect(engine, lib.tp_run(engine.h, C.byref(params), tp.RNG_XOSHIRO, 5, C.byref(st), n, pp, C.byref(merged),
                              None), ERR_ARG, "unknown seeding")
    expect(engine, lib.tp_run(engine.h, C.byref(params), tp.RNG_XOSHIRO128, tp.SEED_SKIPAHEAD, C.byref(st), n, pp,
                              C.byref(merged), None), ERR_ARG, "xoshiro128+")
    bad = tp.TpParams.from_buffer_copy(params)
    bad.e_total = float("nan")
    expect(engine, lib.tp_run(engine.h, C.byref(bad), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), n, pp,
                              C.byref(merged), None), ERR_ARG, "e_total")
    bad = tp.TpParams.from_buffer_copy(params)
    bad.precision = 9
    expect(engine, lib.tp_run(engine.h, C.byref(bad), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), n, pp,
                              C.byref(merged), None), ERR_UNSUPPORTED, "precision")
    rs = tp.rng_state(tp.RNG_RANF)
    rs.index = 56
    expect(engine, lib.tp_run(engine.h, C.byref(params), tp.RNG_RANF, tp.SEED_SKIPAHEAD, C.byref(rs), n, pp,
                              C.byref(merged), None), ERR_ARG, "index out of range")
    assert not per.view(np.uint8).any(), "an error wrote per-batch output"
    assert bytes(st) == words, "an error moved the generator state"
    print("tp_run_device argument errors", flush=True)
    expect(engine, lib.tp_run_device(engine.h, C.byref(params), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), 0,
                                     None, None, None), ERR_ARG, "at least one event")
    expect(engine, lib.tp_run_device(engine.h, C.byref(params), tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, C.byref(st), n,
                                     None, None, None), ERR_ARG, "null per-batch output")
    print(f"valid run after the errors N={n}", flush=True)
    _, again, _ = engine.run(params, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), n)
    assert bytes(again) == bytes(ref)


def test_closure_argument_errors_and_empty_batch(engine, params):
    st = tp.rng_state(tp.RNG_XOSHIRO)
    words = bytes(st)
    out = tp.TpAcc()
    print("tp_simulate_batch argument errors", flush=True)
    expect(engine, lib.tp_simulate_batch(engine.h, C.byref(params), tp.RNG_XOSHIRO, C.byref(st), 640_001,
                                         C.byref(out)), ERR_ARG, "640000")
    expect(engine, lib.tp_simulate_batch(engine.h, C.byref(params), tp.RNG_XOSHIRO, C.byref(st), 10, None),
           ERR_ARG, "null output")
    assert bytes(st) == words
    # a batch of no event: the reference's closure returns an empty accumulator
    # and draws nothing
    out.sigma = 1.0
    assert lib.tp_simulate_batch(engine.h, C.byref(params), tp.RNG_XOSHIRO, C.byref(st), 0, C.byref(out)) == 0
    assert not bytes(out).strip(b"\0") and bytes(st) == words
    print("closure after the errors", flush=True)
    a = engine.simulate_batch(params, tp.RNG_XOSHIRO, tp.rng_state(tp.RNG_XOSHIRO), 10_000)
    _, m, _ = engine.run(params, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO), 10_000)
    assert bytes(a) == bytes(m)


def test_events_and_adjoint_argument_errors(engine, params):
    st = tp.rng_state(tp.RNG_XOSHIRO, 42)
    print("tp_events / tp_adjoint argument errors", flush=True)
    expect(engine, lib.tp_events(engine.h, C.byref(params), tp.RNG_XOSHIRO, C.byref(st), (1 << 20) + 1, None, None,
                                 None, None), ERR_ARG, "2^20")
    F = np.zeros((16, 5))
    expect(engine, lib.tp_adjoint(engine.h, C.byref(params), tp.RNG_XOSHIRO, C.byref(st), None, 16, tp.PREC_MIXED, 0,
                                  F.ctypes.data_as(C.c_void_p), None, None), ERR_ARG, "precision")
    expect(engine, lib.tp_adjoint(engine.h, None, tp.RNG_XOSHIRO, C.byref(st), None, 16, tp.PREC_F64, 0,
                                  F.ctypes.data_as(C.c_void_p), None, None), ERR_ARG, "null params")
    assert not F.any()
    print("adjoint after the errors", flush=True)
    F1, gu1, gp1 = engine.adjoint(params, 1000, st=tp.rng_state(tp.RNG_XOSHIRO, 42))
    F2, gu2, gp2 = engine.adjoint(params, 1000, st=tp.rng_state(tp.RNG_XOSHIRO, 42))
    assert np.array_equal(F1, F2) and np.array_equal(gu1, gu2) and np.array_equal(gp1, gp2)


def test_flow_argument_errors(engine, params):
    f = tp.Flow(engine, init_seed=3)
    try:
        print("tp_flow_train / tp_flow_run argument errors", flush=True)
        for kw, text in (({"batch": 1000}, "multiple of 16"), ({"lr": 0.0}, "must be positive"),
                         ({"target": 3}, "target"), ({"target": 2, "cap_rel": 0.0}, "cap_rel")):
            o = tp.flow_default_opts(steps=1, **kw)
            expect(engine, lib.tp_flow_train(f.h, C.byref(params), C.byref(o), None, None), ERR_ARG, text)
        per = np.zeros(2, dtype=tp.ACC_DTYPE)
        pp = per.ctypes.data_as(C.c_void_p)
        expect(engine, lib.tp_flow_run(f.h, C.byref(params), 1.5, 1, 20_000, pp, None, None), ERR_ARG, "alpha")
        expect(engine, lib.tp_flow_run(f.h, C.byref(params), 0.9, 1, 0, pp, None, None), ERR_ARG,
               "at least one event")
        assert not per.view(np.uint8).any()
        w = f.get_params()
        assert not w[tp.FLOW_LAYER - 32 * 64 - 32:tp.FLOW_LAYER].any(), "failed training calls changed the weights"
        print("weighted run after the errors", flush=True)
        _, a, _ = f.run(params, 0.9, 1, 20_000)
        _, b, _ = f.run(params, 0.9, 1, 20_000)
        assert bytes(a) == bytes(b) and a.selected_events > 0
    finally:
        f.close()
