// tp_adjstore.h -- where the f32 adjoint kernel's staged rows go.
//
// A 64-lane workgroup takes the points base .. base + np - 1 (np = the
// points left, at most 64; lane l computes point base + l when l < np and a
// stand-in point otherwise).  It stages its rows in LDS and stores them with
// consecutive lanes on consecutive addresses:
//   F[n][5]         its 5 np values, contiguous from 5 base;
//   dF/du[n][5][12] its 60 np values, contiguous from 60 base, after the
//                   five sweeps;
//   dF/dp[n][5][12] per sweep c, 12 values of each of its points (element e
//                   of the stage is point e / 12, column e % 12).
// Every element index below is < 5 np, 60 np or 12 np, so every store
// stays inside the rows of the workgroup's own points (< n).  The same
// functions build for the host, where tests/adjstore_check.cpp checks them
// for every lane and every n up to a few workgroups.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TPA_FN __host__ __device__ __forceinline__
#else
#define TPA_FN inline
#endif

namespace tpa {

constexpr uint32_t LANES = 64;

// Points of the workgroup that starts at base (n > base).
TPA_FN uint32_t points(uint64_t n, uint64_t base) { return n - base < LANES ? (uint32_t)(n - base) : LANES; }
// Stage slots of a lane's values (LDS, per workgroup).
TPA_FN uint32_t slot_f(uint32_t lane, int c) { return 5 * lane + (uint32_t)c; }
TPA_FN uint32_t slot_du(uint32_t lane, int c, int j) { return 60 * lane + 12 * (uint32_t)c + (uint32_t)j; }
TPA_FN uint32_t slot_dp(uint32_t lane, int j) { return 12 * lane + (uint32_t)j; }
// Global index of stage element e (e < 5 np, 60 np, 12 np respectively).
TPA_FN uint64_t dst_f(uint64_t base, uint32_t e) { return 5 * base + e; }
TPA_FN uint64_t dst_du(uint64_t base, uint32_t e) { return 60 * base + e; }
TPA_FN uint64_t dst_dp(uint64_t base, uint32_t e, int c) { return 60 * (base + e / 12) + 12 * (uint64_t)c + e % 12; }

} // namespace tpa
