"""Writes f32_conditioning_event.json: the RANF state (default seed) just
before event 6779 of batch 8955 of the f32 faster-evgen stream, an event whose
squared amplitudes f32 arithmetic gets 5-11 % away from f64's (the f32 oracle
against the f64 oracle from the same fifteen draws, none of them re-rolled).

The state comes from the f32 oracle advanced event by event through the
batches before it (sequential.rs:26-36 with the faster-evgen draws,
evgen.rs:110-170); about 20 s of CPU.

    python3 tests/golden/make_f32_conditioning.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_ffi import Oracle  # noqa: E402

BATCH, EVENT = 8955, 6779


def main():
    o = Oracle(f32=True)
    cfg = o.config(os.path.join(HERE, "valeurs"))
    opts = o.opts("ranf", faster_evgen=True)
    g = o.rng("ranf")
    for _ in range(BATCH):
        o.simulate_batch(cfg, opts, g, 10000)
    o.simulate_batch(cfg, opts, g, EVENT)
    out = {"batch": BATCH, "event": EVENT, "ranf_index": int(g.index), "ranf_numbers": [int(x) for x in g.numbers]}
    with open(os.path.join(HERE, "f32_conditioning_event.json"), "w") as f:
        json.dump(out, f)
        f.write("\n")


if __name__ == "__main__":
    main()
