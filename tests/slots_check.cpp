// slots_check.cpp -- TEST INFRASTRUCTURE ONLY: the event loop's sort-slot
// function (3photons-rust_amd/csrc/tp_physics.h, tpd::sort_slots, built here
// for the host) against the reference's literal swap network
// (evgen.rs:109-118: strict '>' swaps over the slot pairs (0,1), (0,2), (1,2))
// on every combination of three energies drawn from a small set that holds
// ties, signed zeros and NaN.
//
//   hipcc -x hip --offload-host-only -O2 -std=c++17 -shared -fPIC -I<csrc> slots_check.cpp
#include "tp_physics.h"

#include <cmath>

namespace {
void network(const double e[3], int d[3])
{
    double v[3] = {e[0], e[1], e[2]};
    int id[3] = {0, 1, 2};
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (const auto &pr : pairs) {
        const int i = pr[0], j = pr[1];
        if (v[j] > v[i]) {
            const double tv = v[i];
            v[i] = v[j];
            v[j] = tv;
            const int ti = id[i];
            id[i] = id[j];
            id[j] = ti;
        }
    }
    for (int slot = 0; slot < 3; ++slot)
        d[id[slot]] = slot;
}
} // namespace

// Returns the number of combinations where the two disagree; *cases gets
// the number of combinations tried.
extern "C" int tp_slots_mismatches(int *cases)
{
    const double vals[] = {0.0, -0.0, 1.0, 2.0, -1.0, 1e300, NAN};
    const int nv = sizeof vals / sizeof vals[0];
    int bad = 0, n = 0;
    for (int a = 0; a < nv; ++a)
        for (int b = 0; b < nv; ++b)
            for (int c = 0; c < nv; ++c) {
                double p[3][4] = {};
                p[0][3] = vals[a];
                p[1][3] = vals[b];
                p[2][3] = vals[c];
                const double e[3] = {vals[a], vals[b], vals[c]};
                int want[3], got[3];
                network(e, want);
                tpd::sort_slots(p, 1, got);
                bad += (want[0] != got[0] || want[1] != got[1] || want[2] != got[2]);
                int id[3];
                tpd::sort_slots(p, 0, id); // sorting off: identity
                bad += (id[0] != 0 || id[1] != 1 || id[2] != 2);
                ++n;
            }
    *cases = n;
    return bad;
}
