// opcount.cpp -- counts the arithmetic the reference performs per generated
// event by instantiating the oracle's per-event code (or_event.h) with a
// counting scalar.  TEST INFRASTRUCTURE ONLY (it fixes the algorithmic work
// figure that the benchmark divides by kernel time; nothing in the product
// uses it).
//
// Counting rules: each +, -, *, /, sqrt and each sin/cos/ln is one operation;
// negation, abs and comparisons are free.  The twelve uniform conversions
// (one multiply each: RANF int * 1e-9, xoshiro (u64 >> 11) * 2^-53) are added.
// Events follow the reference's control flow exactly, including the early
// returns of EventCut::keep (evcut.rs:42-96).
//
// Usage: opcount [events] [ranf|xoshiro]   -> one JSON line on stdout.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {
struct Counts {
    long long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
} C;
} // namespace

struct Cnt {
    double v;
    Cnt(double x = 0.) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { ++C.add; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++C.add; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++C.mul; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++C.div; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt &operator*=(Cnt &a, Cnt b) { a = a * b; return a; }
inline Cnt &operator+=(Cnt &a, Cnt b) { a = a + b; return a; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline Cnt cnt_sqrt(Cnt a) { ++C.sqrt; return Cnt(std::sqrt(a.v)); }
inline Cnt cnt_sin(Cnt a) { ++C.trans; return Cnt(std::sin(a.v)); }
inline Cnt cnt_cos(Cnt a) { ++C.trans; return Cnt(std::cos(a.v)); }
inline Cnt cnt_log(Cnt a) { ++C.trans; return Cnt(std::log(a.v)); }
inline Cnt cnt_fabs(Cnt a) { return Cnt(std::fabs(a.v)); }

#define OR_REAL Cnt
#define OR_SQRT cnt_sqrt
#define OR_SIN cnt_sin
#define OR_COS cnt_cos
#define OR_LOG cnt_log
#define OR_FABS cnt_fabs
#include "or_event.h"

namespace {
uint64_t xs[4];
uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
double xo_uniform()
{
    uint64_t r = xs[0] + xs[3], t = xs[1] << 17;
    xs[2] ^= xs[0];
    xs[3] ^= xs[1];
    xs[1] ^= xs[2];
    xs[0] ^= xs[3];
    xs[2] ^= t;
    xs[3] = rotl(xs[3], 45);
    return (double)(r >> 11) * 0x1.0p-53;
}
uint64_t splitmix(uint64_t &x)
{
    x += 0x9e3779b97f4a7c15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
} // namespace

int main(int argc, char **argv)
{
    long long n = argc > 1 ? std::atoll(argv[1]) : 1000000;
    uint64_t seed = 12345;
    for (int i = 0; i < 4; ++i)
        xs[i] = splitmix(seed);
    // Shipped valeurs (valeurs:1-18) and the constants the reference derives
    // from them; the operation counts do not depend on their exact values.
    or_phys ph;
    ph.e_total = 91.187;
    ph.cut_beam = 0.9;
    ph.cut_pp = 0.9396;
    ph.e_min = 4.559;
    ph.cut_plane = 0.;
    ph.g_a = -0.0270890588;
    ph.g_bp = ph.g_bm = -1.2e-8;
    for (int c = 0; c < 5; ++c)
        ph.sc[c] = 1e-3 * (c + 1);
    ph.sort = 1;
    long long kept = 0;
    Counts before_gen, after_gen, after_cut;
    for (long long e = 0; e < n; ++e) {
        Cnt u[12];
        for (int j = 0; j < 12; ++j) {
            u[j] = Cnt(xo_uniform());
            ++C.mul; // the int -> [0,1) scaling multiply
        }
        Cnt p[5][4];
        or_generate(&ph, u, p);
        if (!or_keep(&ph, p))
            continue;
        ++kept;
        Cnt m2[5][8], m[5];
        or_m2(&ph, p, m2);
        or_m2_sums(m2, m);
        // ResultsAccumulator::integrate (resacc.rs:121-129).
        Cnt spm2[5], vars[5], sigma, var;
        for (int c = 0; c < 5; ++c)
            spm2[c] = spm2[c] + m[c];
        for (int c = 0; c < 5; ++c)
            vars[c] = vars[c] + m[c] * m[c];
        Cnt w = or_weight(&ph, m);
        sigma = sigma + w;
        var = var + w * w;
    }
    double N = (double)n;
    double total = (C.add + C.mul + C.div + C.sqrt + C.trans) / N;
    std::printf("{\"events\": %lld, \"pass_rate\": %.7f, \"add_sub\": %.4f, \"mul\": %.4f, \"div\": %.4f, "
                "\"sqrt\": %.4f, \"transcendental\": %.4f, \"ops_per_event\": %.4f}\n",
                n, kept / N, C.add / N, C.mul / N, C.div / N, C.sqrt / N, C.trans / N, total);
    return 0;
}
