"""tp_run_device_noalloc (no device allocation; tables in a static device
array) against tp_run_device: the same per-batch accumulators and in-order
fold, bit for bit, on two calls in a row, the second with a smaller input,
for both precisions and both seedings; and its clean refusals."""
import ctypes as C

import numpy as np
import pytest

import trois_photons as tp

pytestmark = pytest.mark.gpu


def run_both(engine, params, seeding, n, per_a, mer_a, per_b, mer_b, torch):
    nb = (n + tp.BATCH_SIZE - 1) // tp.BATCH_SIZE
    for buf in (per_a, mer_a, per_b, mer_b):
        buf.fill_(0xA5)  # stale bytes: every accumulator must be written
    engine.run_device(params, tp.RNG_XOSHIRO, seeding, tp.rng_state(tp.RNG_XOSHIRO), n, per_a.data_ptr(),
                      mer_a.data_ptr())
    engine.sync()
    engine.run_device_noalloc(params, tp.RNG_XOSHIRO, seeding, tp.rng_state(tp.RNG_XOSHIRO), n, per_b.data_ptr(),
                              mer_b.data_ptr())
    engine.sync()
    a = per_a[:nb * 104].cpu().numpy().tobytes()
    b = per_b[:nb * 104].cpu().numpy().tobytes()
    assert a == b, f"per-batch accumulators differ (n={n})"
    assert mer_a.cpu().numpy().tobytes() == mer_b.cpu().numpy().tobytes(), f"fold differs (n={n})"
    acc = tp.TpAcc.from_buffer_copy(mer_b.cpu().numpy().tobytes())
    assert acc.selected_events > 0.6 * n


@pytest.mark.parametrize("prec,seeding", [(tp.PREC_F64, tp.SEED_SKIPAHEAD), (tp.PREC_F64, tp.SEED_JUMP),
                                          (tp.PREC_MIXED, tp.SEED_SKIPAHEAD)])
def test_noalloc_matches_run_device(valeurs, prec, seeding):
    import torch
    cfg = tp.Configuration.load(valeurs)
    dev = torch.device("cuda:0")
    n1, n2 = 40_000_000, 123_457  # 4000 batches (past one pass of the persistent grid), then a partial batch
    per_a = torch.empty(4000 * 104, dtype=torch.uint8, device=dev)
    per_b = torch.empty_like(per_a)
    mer_a = torch.empty(104, dtype=torch.uint8, device=dev)
    mer_b = torch.empty_like(mer_a)
    with tp.Engine(0) as engine:
        for n in (n1, n2):
            print(f"noalloc prec={prec} seeding={seeding} n={n}", flush=True)
            run_both(engine, tp.params_from_config(cfg.with_events(n), precision=prec), seeding, n,
                     per_a, mer_a, per_b, mer_b, torch)


def test_noalloc_refuses_what_it_does_not_run(valeurs):
    import torch
    cfg = tp.Configuration.load(valeurs)
    buf = torch.empty(104, dtype=torch.uint8, device=torch.device("cuda:0"))
    with tp.Engine(0) as engine:
        p = tp.params_from_config(cfg.with_events(10))
        with pytest.raises(tp.TpError, match="xoshiro256"):
            engine.run_device_noalloc(p, tp.RNG_RANF, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_RANF), 10,
                                      buf.data_ptr(), None)
        with pytest.raises(tp.TpError, match="2\\^32"):
            engine.run_device_noalloc(p, tp.RNG_XOSHIRO, tp.SEED_SKIPAHEAD, tp.rng_state(tp.RNG_XOSHIRO),
                                      (1 << 32) * tp.BATCH_SIZE, buf.data_ptr(), None)
