// walk_check.cpp -- TEST INFRASTRUCTURE ONLY: the faster-evgen stream walk
// (3photons-rust_amd/csrc/tp_rng_host.cpp) split over threads against the
// same walk done sequentially: event starts, end position and final state;
// and the walks' end positions, which tests/test_walk.py checks against the
// oracle's closure.
//
//   g++ -O2 -std=c++17 -pthread -shared -fPIC -I<csrc> walk_check.cpp <csrc>/tp_rng_host.cpp
#include "tp_rng_host.h"

#include <cstring>
#include <vector>

// 0 when the threaded walk of n events from seed / index equals the
// sequential one.
extern "C" int tp_walk_mismatch(uint64_t n, int seed, uint32_t index, int threads, int f32)
{
    tp::RanfState g0;
    tp::ranf_seeded(g0, seed);
    g0.index = index;
    tp::RanfState a = g0, b = g0;
    std::vector<uint64_t> sa(n), sb(n);
    uint64_t qa = 0, qb = 0;
    tp::ranf_walk_faster_evgen(a, n, sa.data(), &qa, 1, f32 != 0);
    tp::ranf_walk_faster_evgen(b, n, sb.data(), &qb, threads, f32 != 0);
    int bad = 0;
    bad += qa != qb;
    bad += sa != sb;
    bad += a.index != b.index || std::memcmp(a.numbers, b.numbers, sizeof a.numbers) != 0;
    return bad;
}

// 0 when the walk of n events from the start of event e0 (the way a chunk of
// a long run is walked) equals events e0 .. e0 + n of the walk from the
// state itself: starts and end position.
extern "C" int tp_walk_from_mismatch(uint64_t e0, uint64_t n, int seed, uint32_t index, int threads)
{
    tp::RanfState g0;
    tp::ranf_seeded(g0, seed);
    g0.index = index;
    tp::RanfState a = g0;
    std::vector<uint64_t> sa(e0 + n), sb(n);
    uint64_t qa = 0, qb = 0;
    tp::ranf_walk_faster_evgen(a, e0 + n, sa.data(), &qa, 1);
    const uint64_t q = e0 < e0 + n ? sa[e0] : qa;
    tp::ranf_walk_faster_evgen_from(g0, q, n, sb.data(), &qb, threads);
    int bad = qa != qb;
    for (uint64_t e = 0; e < n; ++e)
        bad += sa[e0 + e] != sb[e];
    return bad;
}

// Stream position, counted from the start of the round of the state
// (numbers[1..55], index), where event n of the faster-evgen walk from that
// state starts: the position a batch beginning with event n starts from.
extern "C" uint64_t tp_walk_end(const int32_t *numbers, uint32_t index, uint64_t n, int threads, int f32)
{
    tp::RanfState g;
    std::memcpy(g.numbers, numbers, sizeof g.numbers);
    g.index = index;
    uint64_t q = 0;
    tp::ranf_walk_faster_evgen(g, n, nullptr, &q, threads, f32 != 0);
    return q;
}

// Draws after s0 where event n of the faster-evgen walk of the xoshiro256+
// stream from s0 starts.
extern "C" uint64_t tp_xo_walk_end(const uint64_t *s0, uint64_t n)
{
    uint64_t q = 0;
    tp::xoshiro_walk_faster_evgen(s0, 0, n, nullptr, &q);
    return q;
}

// The same on the xoshiro128+ stream of the all-f32 mode.
extern "C" uint64_t tp_xo128_walk_end(const uint32_t *s0, uint64_t n)
{
    uint64_t q = 0;
    tp::xoshiro128_walk_faster_evgen(s0, 0, n, nullptr, &q);
    return q;
}
