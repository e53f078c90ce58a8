// tp_rng_host.cpp -- host-side stream algebra (see tp_rng_host.h).
#include "tp_rng_host.h"

#include <algorithm>
#include <cstring>
#include <system_error>
#include <thread>
#include <limits>
#include <vector>
#include <mutex>

namespace tp {

// ---- RANF ------------------------------------------------------------------

// The reference's two loops, with x[i] += 1e9 where the difference is
// negative done without a branch; the second loop is cut at i = 49 so that
// no block reads what it writes (x[i - 24] for i < 49 comes from the first
// loop, for i >= 49 from the first part of the second), which lets the
// compiler vectorise each block.
void ranf_round(int32_t *x)
{
    for (int i = 1; i < 25; ++i) {
        const int32_t v = x[i] - x[i + 31];
        x[i] = v + ((v >> 31) & RANF_MODULO);
    }
    for (int i = 25; i < 49; ++i) {
        const int32_t v = x[i] - x[i - 24];
        x[i] = v + ((v >> 31) & RANF_MODULO);
    }
    for (int i = 49; i < 56; ++i) {
        const int32_t v = x[i] - x[i - 24];
        x[i] = v + ((v >> 31) & RANF_MODULO);
    }
}

void ranf_seeded(RanfState &g, int32_t seed)
{
    std::memset(&g, 0, sizeof g);
    g.index = 55;
    g.seed = seed;
    g.numbers[55] = seed;
    int32_t j = seed, k = 1;
    for (int i = 1; i < 55; ++i) {
        int ii = (21 * i) % 55;
        g.numbers[ii] = k;
        // i32 arithmetic wraps in the reference's release build.
        k = static_cast<int32_t>(static_cast<uint32_t>(j) - static_cast<uint32_t>(k));
        if (k < 0)
            k = static_cast<int32_t>(static_cast<uint32_t>(k) + static_cast<uint32_t>(RANF_MODULO));
        j = g.numbers[ii];
    }
    for (int r = 0; r < 10; ++r)
        ranf_round(g.numbers);
}

static inline uint64_t mulmod(uint64_t a, uint64_t b) { return (a * b) % RANF_MODULO; }

RanfPoly ranf_poly_mul(const RanfPoly &a, const RanfPoly &b)
{
    // Product of degree <= 108, then reduce with x^55 = 1 - x^31.
    uint64_t prod[109] = {0};
    for (int i = 0; i < 55; ++i) {
        if (!a[i])
            continue;
        for (int j = 0; j < 55; ++j)
            prod[i + j] = (prod[i + j] + mulmod(a[i], b[j])) % RANF_MODULO;
    }
    for (int d = 108; d >= 55; --d) {
        uint64_t c = prod[d];
        if (!c)
            continue;
        prod[d] = 0;
        prod[d - 55] = (prod[d - 55] + c) % RANF_MODULO;
        prod[d - 24] = (prod[d - 24] + RANF_MODULO - c) % RANF_MODULO;
    }
    RanfPoly r(55);
    for (int i = 0; i < 55; ++i)
        r[i] = static_cast<uint32_t>(prod[i]);
    return r;
}

RanfPoly ranf_poly_xpow(uint64_t n)
{
    RanfPoly result(55, 0), base(55, 0);
    result[0] = 1;
    // base = x (or reduced higher power once n is large)
    base[1] = 1;
    while (n) {
        if (n & 1)
            result = ranf_poly_mul(result, base);
        n >>= 1;
        if (n)
            base = ranf_poly_mul(base, base);
    }
    return result;
}

void ranf_apply_round_jump(const RanfPoly &c, const int32_t *x_in, int32_t *x_out)
{
    // Window Y[0..109] = X_r[1..55], X_{r+1}[1..55]; X_{r+m}[i] = sum_k c_k Y[i-1+k].
    int32_t next[56];
    std::memcpy(next, x_in, sizeof next);
    ranf_round(next);
    uint64_t y[110];
    for (int i = 0; i < 55; ++i) {
        y[i] = static_cast<uint64_t>(x_in[i + 1]);
        y[55 + i] = static_cast<uint64_t>(next[i + 1]);
    }
    x_out[0] = 0;
    for (int i = 0; i < 55; ++i) {
        uint64_t acc = 0;
        for (int k = 0; k < 55; ++k)
            acc = (acc + mulmod(c[k], y[i + k])) % RANF_MODULO;
        x_out[i + 1] = static_cast<int32_t>(acc);
    }
}

void ranf_advance(RanfState &g, uint64_t n)
{
    if (n <= g.index) {
        g.index -= static_cast<uint32_t>(n);
        return;
    }
    uint64_t rest = n - g.index;
    uint64_t rounds = (rest + 54) / 55;
    RanfPoly c = ranf_poly_xpow(55 * rounds);
    int32_t out[56];
    ranf_apply_round_jump(c, g.numbers, out);
    std::memcpy(g.numbers, out, sizeof out);
    g.index = static_cast<uint32_t>(55 * rounds - rest);
}

// ---- xoshiro256+ -------------------------------------------------------------

static uint64_t splitmix64(uint64_t &x)
{
    x += 0x9e3779b97f4a7c15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void xoshiro_seed(uint64_t seed, uint64_t s[4])
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i)
        s[i] = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline void xo_update(uint64_t *s)
{
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
}

uint64_t xoshiro_next(uint64_t s[4])
{
    uint64_t r = s[0] + s[3];
    xo_update(s);
    return r;
}

void gf2_identity(Gf2Mat &m)
{
    std::memset(&m, 0, sizeof m);
    for (int i = 0; i < 256; ++i)
        m.col[i][i / 64] = 1ull << (i % 64);
}

void gf2_step(Gf2Mat &m)
{
    gf2_identity(m);
    for (int i = 0; i < 256; ++i)
        xo_update(m.col[i]);
}

void gf2_jump128(Gf2Mat &m)
{
    static const uint64_t J[4] = {0x180ec6d33cfd0abaull, 0xd5a61266f0c9392cull, 0xa9582618e03fc9aaull,
                                  0x39abdc4529b1661cull};
    for (int i = 0; i < 256; ++i) {
        uint64_t s[4] = {0, 0, 0, 0};
        s[i / 64] = 1ull << (i % 64);
        uint64_t a[4] = {0, 0, 0, 0};
        for (int w = 0; w < 4; ++w)
            for (int b = 0; b < 64; ++b) {
                if (J[w] & (1ull << b))
                    for (int q = 0; q < 4; ++q)
                        a[q] ^= s[q];
                xo_update(s);
            }
        std::memcpy(m.col[i], a, sizeof a);
    }
}

void gf2_apply(const Gf2Mat &m, const uint64_t in[4], uint64_t out[4])
{
    uint64_t r[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w) {
        uint64_t x = in[w];
        while (x) {
            int b = __builtin_ctzll(x);
            x &= x - 1;
            const uint64_t *c = m.col[w * 64 + b];
            r[0] ^= c[0];
            r[1] ^= c[1];
            r[2] ^= c[2];
            r[3] ^= c[3];
        }
    }
    std::memcpy(out, r, sizeof r);
}

void gf2_mul(const Gf2Mat &a, const Gf2Mat &b, Gf2Mat &out)
{
    Gf2Mat tmp;
    for (int i = 0; i < 256; ++i)
        gf2_apply(a, b.col[i], tmp.col[i]);
    out = tmp;
}

std::vector<Gf2Mat> gf2_pow2_table(const Gf2Mat &base, int count)
{
    std::vector<Gf2Mat> t(static_cast<size_t>(count));
    if (count == 0)
        return t;
    t[0] = base;
    for (int k = 1; k < count; ++k)
        gf2_mul(t[k - 1], t[k - 1], t[k]);
    return t;
}

void gf2_pow_from_table(const std::vector<Gf2Mat> &table, uint64_t n, Gf2Mat &out)
{
    gf2_identity(out);
    for (size_t k = 0; n && k < table.size(); ++k, n >>= 1)
        if (n & 1)
            gf2_mul(table[k], out, out);
}

const std::vector<Gf2Mat> &xoshiro_step_pow2()
{
    static std::once_flag once;
    static std::vector<Gf2Mat> table;
    std::call_once(once, [] {
        Gf2Mat m;
        gf2_step(m);
        table = gf2_pow2_table(m, 64);
    });
    return table;
}

void xoshiro_advance(uint64_t s[4], uint64_t n)
{
    const std::vector<Gf2Mat> &t = xoshiro_step_pow2();
    for (int k = 0; n; ++k, n >>= 1)
        if (n & 1)
            gf2_apply(t[static_cast<size_t>(k)], s, s);
}

// ---- xoshiro128+ ---------------------------------------------------------------

void xoshiro128_seed(uint64_t seed, uint32_t s[4])
{
    uint64_t x = seed;
    const uint64_t a = splitmix64(x), b = splitmix64(x);
    s[0] = (uint32_t)a;
    s[1] = (uint32_t)(a >> 32);
    s[2] = (uint32_t)b;
    s[3] = (uint32_t)(b >> 32);
}

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

static inline void xo128_update(uint32_t *s)
{
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
}

uint32_t xoshiro128_next(uint32_t s[4])
{
    const uint32_t r = s[0] + s[3];
    xo128_update(s);
    return r;
}

static const uint32_t XO128_JUMP[4] = {0x8764000bu, 0xf542d2d3u, 0x6fa035c3u, 0x77f2db5bu};

static void xo128_jump_state(uint32_t *s)
{
    uint32_t a[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w)
        for (int b = 0; b < 32; ++b) {
            if (XO128_JUMP[w] & (1u << b))
                for (int q = 0; q < 4; ++q)
                    a[q] ^= s[q];
            xo128_update(s);
        }
    std::memcpy(s, a, sizeof a);
}

void xoshiro128_jump(uint32_t s[4]) { xo128_jump_state(s); }

void gf2_128_identity(Gf2Mat128 &m)
{
    std::memset(&m, 0, sizeof m);
    for (int i = 0; i < 128; ++i)
        m.col[i][i / 32] = 1u << (i % 32);
}

void gf2_128_step(Gf2Mat128 &m)
{
    gf2_128_identity(m);
    for (int i = 0; i < 128; ++i)
        xo128_update(m.col[i]);
}

void gf2_128_jump64(Gf2Mat128 &m)
{
    gf2_128_identity(m);
    for (int i = 0; i < 128; ++i)
        xo128_jump_state(m.col[i]);
}

void gf2_128_apply(const Gf2Mat128 &m, const uint32_t in[4], uint32_t out[4])
{
    uint32_t r[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w) {
        uint32_t x = in[w];
        while (x) {
            const int b = __builtin_ctz(x);
            x &= x - 1;
            const uint32_t *c = m.col[w * 32 + b];
            r[0] ^= c[0];
            r[1] ^= c[1];
            r[2] ^= c[2];
            r[3] ^= c[3];
        }
    }
    std::memcpy(out, r, sizeof r);
}

void gf2_128_mul(const Gf2Mat128 &a, const Gf2Mat128 &b, Gf2Mat128 &out)
{
    Gf2Mat128 tmp;
    for (int i = 0; i < 128; ++i)
        gf2_128_apply(a, b.col[i], tmp.col[i]);
    out = tmp;
}

std::vector<Gf2Mat128> gf2_128_pow2_table(const Gf2Mat128 &base, int count)
{
    std::vector<Gf2Mat128> t(static_cast<size_t>(count));
    if (count == 0)
        return t;
    t[0] = base;
    for (int k = 1; k < count; ++k)
        gf2_128_mul(t[k - 1], t[k - 1], t[k]);
    return t;
}

void gf2_128_pow_from_table(const std::vector<Gf2Mat128> &table, uint64_t n, Gf2Mat128 &out)
{
    gf2_128_identity(out);
    for (size_t k = 0; n && k < table.size(); ++k, n >>= 1)
        if (n & 1)
            gf2_128_mul(table[k], out, out);
}

const std::vector<Gf2Mat128> &xoshiro128_step_pow2()
{
    static std::once_flag once;
    static std::vector<Gf2Mat128> table;
    std::call_once(once, [] {
        Gf2Mat128 m;
        gf2_128_step(m);
        table = gf2_128_pow2_table(m, 64);
    });
    return table;
}

void xoshiro128_advance(uint32_t s[4], uint64_t n)
{
    const std::vector<Gf2Mat128> &t = xoshiro128_step_pow2();
    for (int k = 0; n; ++k, n >>= 1)
        if (n & 1)
            gf2_128_apply(t[static_cast<size_t>(k)], s, s);
}

// ---- faster-evgen stream walk ----------------------------------------------
namespace {
struct RanfWalker {
    RanfState g;
    uint64_t q; // position of the next value
    // random_array::<N>: the N values in the order the reference returns them
    // (ascending array index = descending stream position), as raw integers.
    const int32_t *take(int n)
    {
        if ((int)g.index < n) {
            q += g.index; // leftovers of the round are discarded
            ranf_round(g.numbers);
            g.index = 55;
        }
        g.index -= (uint32_t)n;
        q += (uint64_t)n;
        return g.numbers + g.index + 1;
    }
};
} // namespace

namespace {
// Event starts from the walker's position onward, until n events are recorded
// or the position reaches q_stop (when q_stop is not 0).  T is the run's float
// type: the unit-disc test runs on (n as T) * 1e-9 values (ranf.rs:99), and
// f32's MIN_POSITIVE squared underflows to 0, as in the reference's f32 build.
template <class T> void walk_events_t(RanfWalker &w, uint64_t n, uint64_t q_stop, std::vector<uint64_t> &out)
{
    const T minpos = std::numeric_limits<T>::min();
    const T min2 = minpos * minpos;
    for (uint64_t e = 0; e < n; ++e) {
        if (q_stop && w.q >= q_stop)
            return;
        out.push_back(w.q);
        w.take(9); // only their positions matter here
        const int32_t *r6 = w.take(6);
        T px[3], py[3];
        for (int i = 0; i < 3; ++i) {
            px[i] = (T)r6[i] * (T)1e-9;
            py[i] = (T)r6[3 + i] * (T)1e-9;
        }
        // radii of the three points (evgen.rs:228-246); only the accept test
        // matters here, with the reference's float operations
        for (int i = 0; i < 3; ++i) {
            const T x = (T)2 * px[i] - (T)1, y = (T)2 * py[i] - (T)1;
            T r2 = ((T)0 + x * x) + y * y;
            while (r2 > (T)1 || r2 < min2) {
                const int32_t *r2v = w.take(2);
                const T a = (T)2 * ((T)r2v[0] * (T)1e-9) - (T)1, b = (T)2 * ((T)r2v[1] * (T)1e-9) - (T)1;
                r2 = (T)0 + (a * a + b * b);
            }
        }
    }
}
void walk_events(RanfWalker &w, uint64_t n, uint64_t q_stop, std::vector<uint64_t> &out, bool f32)
{
    if (f32)
        walk_events_t<float>(w, n, q_stop, out);
    else
        walk_events_t<double>(w, n, q_stop, out);
}

// A walker whose next value is at stream position q (q >= q0, the position
// of g0's next value).
RanfWalker walker_at(const RanfState &g0, uint64_t q0, uint64_t q)
{
    RanfState g = g0;
    ranf_advance(g, q - q0);
    return RanfWalker{g, q};
}
} // namespace

// The walk is sequential by nature, but a walker's future depends only on its
// stream position at an event start.  Long walks therefore split the stream
// at round boundaries: each thread starts at its boundary as if an event
// started there, and walks its segment plus an overlap into the next.  The
// true walk from the previous segment reaches, within a few events, an event
// start the next thread also recorded; from there on both agree, so the
// segments are joined at that point (a sequential walk covers any gap).
void ranf_walk_faster_evgen(RanfState &g, uint64_t num_events, uint64_t *starts, uint64_t *q_end, int max_threads,
                            bool f32)
{
    const uint64_t q0 = 55u - g.index;
    unsigned hw = std::thread::hardware_concurrency();
    const int T = (int)std::min<unsigned>((unsigned)std::max(1, max_threads), hw ? hw : 1u);
    // The true walk as pieces [begin, end) of the lists, in stream order.
    struct Piece {
        size_t list, begin, end;
    };
    std::vector<std::vector<uint64_t>> lists;
    std::vector<Piece> pieces;
    auto sequential = [&]() {
        lists.assign(1, {});
        RanfWalker w{g, q0};
        lists[0].reserve(num_events + 1);
        walk_events(w, num_events + 1, 0, lists[0], f32);
        pieces.assign(1, Piece{0, 0, lists[0].size()});
    };
    // ~17.9 draws per event; T segments of whole rounds, each walked 64 rounds past its end
    const uint64_t est = q0 + (uint64_t)((double)num_events * 18.5);
    const uint64_t seg = (est / (uint64_t)T / 55 + 1) * 55;
    const uint64_t overlap = 64 * 55;
    if (num_events < 4000000 || T < 2) {
        sequential();
    } else {
        lists.resize((size_t)T);
        std::vector<std::thread> th;
        bool started = true;
        try {
            for (int k = 0; k < T; ++k)
                th.emplace_back([&, k] {
                    const uint64_t qs = k == 0 ? q0 : (uint64_t)k * seg;
                    RanfWalker w = k == 0 ? RanfWalker{g, q0} : walker_at(g, q0, qs);
                    lists[(size_t)k].reserve((size_t)(seg / 17));
                    walk_events(w, ~0ull, k + 1 < T ? (uint64_t)(k + 1) * seg + overlap : est + overlap,
                                lists[(size_t)k], f32);
                });
        } catch (const std::system_error &) { // no thread to be had: walk alone
            started = false;
        }
        for (auto &t : th)
            t.join();
        if (!started)
            sequential();
    }
    if (pieces.empty()) {
        // The true walk continues in lists[cur] from index cur_begin; find where
        // it meets list k's speculative starts past that segment's boundary.
        size_t cur = 0, cur_begin = 0;
        uint64_t have = 0;
        for (int k = 1; k < T && have <= num_events; ++k) {
            std::vector<uint64_t> &a = lists[cur];
            const std::vector<uint64_t> &b = lists[(size_t)k];
            const uint64_t qs = (uint64_t)k * seg;
            size_t i = 0, j = 0;
            auto meet = [&]() {
                i = (size_t)(std::lower_bound(a.begin() + (std::ptrdiff_t)cur_begin, a.end(), qs) - a.begin());
                j = 0;
                while (i < a.size() && j < b.size()) {
                    if (a[i] == b[j])
                        return true;
                    if (a[i] < b[j])
                        ++i;
                    else
                        ++j;
                }
                return false;
            };
            bool joined = meet();
            if (!joined && !b.empty()) { // extend the true walk until it meets b or passes it
                RanfWalker w = walker_at(g, q0, a.back());
                uint64_t stop = a.back();
                bool first = true;
                while (!joined && a.back() <= b.back()) {
                    stop += overlap;
                    std::vector<uint64_t> ext;
                    walk_events(w, ~0ull, stop, ext, f32);
                    a.insert(a.end(), ext.begin() + (first ? 1 : 0), ext.end()); // the first repeats a.back()
                    first = false;
                    joined = meet();
                }
            }
            if (!joined) // the true walk has covered this segment itself
                continue;
            pieces.push_back({cur, cur_begin, i});
            have += i - cur_begin;
            cur = (size_t)k;
            cur_begin = j;
        }
        pieces.push_back({cur, cur_begin, lists[cur].size()});
    }
    uint64_t total = 0;
    for (const Piece &p : pieces)
        total += p.end - p.begin;
    if (total < num_events + 1) { // estimate too short: finish sequentially
        const Piece &p = pieces.back();
        RanfWalker w = walker_at(g, q0, lists[p.list][p.end - 1]);
        lists.emplace_back();
        walk_events(w, num_events + 1 - total + 1, 0, lists.back(), f32);
        pieces.push_back({lists.size() - 1, 1, lists.back().size()}); // [0] repeats the previous start
    }
    uint64_t qe = 0, off = 0; // the start of event num_events is where the last one ended
    for (const Piece &p : pieces) {
        const uint64_t *d = lists[p.list].data() + p.begin;
        const uint64_t cnt = p.end - p.begin;
        if (off <= num_events && num_events < off + cnt)
            qe = d[num_events - off];
        if (starts && off < num_events)
            std::memcpy(starts + off, d, (size_t)std::min<uint64_t>(cnt, num_events - off) * sizeof(uint64_t));
        off += cnt;
    }
    ranf_advance(g, qe - q0);
    if (q_end)
        *q_end = qe;
}

void ranf_walk_faster_evgen_from(const RanfState &g0, uint64_t q, uint64_t num_events, uint64_t *starts,
                                 uint64_t *q_end, int max_threads, bool f32)
{
    RanfState g = g0;
    ranf_advance(g, q - (55u - g0.index));
    const uint64_t rel = q - (55u - g.index); // g0-frame position of g's round start
    uint64_t qe = 0;
    ranf_walk_faster_evgen(g, num_events, starts, &qe, max_threads, f32);
    if (starts)
        for (uint64_t e = 0; e < num_events; ++e)
            starts[e] += rel;
    if (q_end)
        *q_end = qe + rel;
}

void xoshiro_walk_faster_evgen(const uint64_t s0[4], uint64_t q, uint64_t num_events, uint64_t *starts,
                               uint64_t *q_end)
{
    uint64_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
    xoshiro_advance(s, q);
    auto uniform = [&]() { return (double)(xoshiro_next(s) >> 11) * 0x1.0p-53; };
    const double min2 = std::numeric_limits<double>::min() * std::numeric_limits<double>::min();
    for (uint64_t e = 0; e < num_events; ++e) {
        if (starts)
            starts[e] = q;
        for (int k = 0; k < 9; ++k) // random_array::<9>: only their positions matter here
            (void)xoshiro_next(s);
        double r6[6];
        for (int k = 0; k < 6; ++k)
            r6[k] = uniform();
        q += 15;
        for (int i = 0; i < 3; ++i) { // evgen.rs:228-246, the reference's float operations
            const double x = 2. * r6[i] - 1., y = 2. * r6[3 + i] - 1.;
            double r2 = (0. + x * x) + y * y;
            while (r2 > 1. || r2 < min2) {
                const double a = 2. * uniform() - 1., b = 2. * uniform() - 1.;
                q += 2;
                r2 = 0. + (a * a + b * b);
            }
        }
    }
    if (q_end)
        *q_end = q;
}

void xoshiro128_walk_faster_evgen(const uint32_t s0[4], uint64_t q, uint64_t num_events, uint64_t *starts,
                                  uint64_t *q_end)
{
    uint32_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
    xoshiro128_advance(s, q);
    auto uniform = [&]() { return (float)(xoshiro128_next(s) >> 8) * 0x1.0p-24f; };
    for (uint64_t e = 0; e < num_events; ++e) {
        if (starts)
            starts[e] = q;
        for (int k = 0; k < 9; ++k) // random_array::<9>: only their positions matter here
            (void)xoshiro128_next(s);
        float r6[6];
        for (int k = 0; k < 6; ++k)
            r6[k] = uniform();
        q += 15;
        for (int i = 0; i < 3; ++i) { // evgen.rs:228-246 in f32: MIN_POSITIVE^2 is 0, only r2 > 1 re-rolls
            const float x = 2.f * r6[i] - 1.f, y = 2.f * r6[3 + i] - 1.f;
            float r2 = (0.f + x * x) + y * y;
            while (r2 > 1.f) {
                const float a = 2.f * uniform() - 1.f, b = 2.f * uniform() - 1.f;
                q += 2;
                r2 = 0.f + (a * a + b * b);
            }
        }
    }
    if (q_end)
        *q_end = q;
}

} // namespace tp
