// Host build of the f32 adjoint kernel's staged-store index arithmetic
// (3photons-rust_amd/csrc/tp_adjstore.h): for every n and every workgroup,
// every lane's stage slots stay inside the stage and every stored element
// lands inside the rows of the workgroup's own points, each written once.
#include "tp_adjstore.h"

#include <cstdio>
#include <vector>

int main()
{
    using namespace tpa;
    for (uint64_t n = 1; n <= 5 * LANES + 7; ++n) {
        std::vector<int> wf(5 * n, 0), wu(60 * n, 0), wp(60 * n, 0);
        for (uint64_t base = 0; base < n; base += LANES) {
            const uint32_t np = points(n, base);
            if (np == 0 || np > LANES || base + np > n)
                return std::printf("n %llu base %llu: np %u\n", (unsigned long long)n, (unsigned long long)base, np), 1;
            for (uint32_t lane = 0; lane < LANES; ++lane) // every lane stages, past np too
                for (int c = 0; c < 5; ++c) {
                    if (slot_f(lane, c) >= 5 * LANES)
                        return std::printf("slot_f\n"), 1;
                    for (int j = 0; j < 12; ++j)
                        if (slot_du(lane, c, j) >= 60 * LANES || slot_dp(lane, j) >= 12 * LANES)
                            return std::printf("slot\n"), 1;
                }
            for (uint32_t e = 0; e < 5 * np; ++e) {
                const uint64_t d = dst_f(base, e);
                if (d < 5 * base || d >= 5 * (base + np) || d >= 5 * n)
                    return std::printf("F out of rows: n %llu\n", (unsigned long long)n), 1;
                ++wf[d];
            }
            for (uint32_t e = 0; e < 60 * np; ++e) {
                const uint64_t d = dst_du(base, e);
                if (d < 60 * base || d >= 60 * (base + np))
                    return std::printf("du out of rows: n %llu\n", (unsigned long long)n), 1;
                ++wu[d];
            }
            for (int c = 0; c < 5; ++c)
                for (uint32_t e = 0; e < 12 * np; ++e) {
                    const uint64_t d = dst_dp(base, e, c);
                    if (d < 60 * base || d >= 60 * (base + np))
                        return std::printf("dp out of rows: n %llu\n", (unsigned long long)n), 1;
                    ++wp[d];
                }
        }
        for (int x : wf)
            if (x != 1)
                return std::printf("F not written once: n %llu\n", (unsigned long long)n), 1;
        for (size_t i = 0; i < wu.size(); ++i)
            if (wu[i] != 1 || wp[i] != 1)
                return std::printf("grad not written once: n %llu\n", (unsigned long long)n), 1;
    }
    std::puts("ok");
    return 0;
}
