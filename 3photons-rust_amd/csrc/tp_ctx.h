// tp_ctx.h -- the engine context shared by the event-loop and flow entry
// points (internal; not part of the C ABI).
#pragma once

#include "tp_guard.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "tp.h"
#include "tp_physics.h"
#include "tp_rng_host.h"

namespace tpe {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

// Temporary device buffers of one call, freed on every return path.
template <int N> struct ScopedBufs {
    DevBuf b[N];
    ScopedBufs() = default;
    ScopedBufs(const ScopedBufs &) = delete;
    ScopedBufs &operator=(const ScopedBufs &) = delete;
    DevBuf &operator[](int i) { return b[i]; }
    ~ScopedBufs()
    {
        for (DevBuf &x : b)
            if (x.p)
                (void)hipFree(x.p);
    }
};

struct Tables { // device copies of GF(2) jump tables for one (seeding, batch size, grid)
    DevBuf bt, stride;
    int kb = 0;
};

struct Ctx {
    int device = 0;
    int cus = 1;                    // compute units of the device
    int blocks_per_cu[5][2] = {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}; // resident fused workgroups per CU [source][precision]
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr}; // [4]: the device walk's status copy
    std::string err;
    DevBuf mt;       // M^(2^k)
    int km = 0;
    std::map<std::string, std::unique_ptr<Tables>> tables;
    std::map<std::string, std::vector<uint32_t>> ranf_polys; // host copies keyed by R
    DevBuf polys, draws, origin, p0, per_batch, merged, evq;
    DevBuf mt128, bt128, f32ev; // all-f32 mode: xoshiro128+ tables, per-event values
    std::string bt128_key;
    int bt128_kb = 0;
    DevBuf st128;               // all-f32 mode: B^grid, the grid-stride advance
    std::string st128_key;
    tp_stats last{};
    bool timed = false;
    // Host staging of the RANF origin: must outlive the asynchronous copies.
    int32_t origin0[56];
    int32_t p00 = 0;
    // faster-evgen: event start positions of the last run and the state after it
    std::vector<uint64_t> fe_starts;
    tp::RanfState fe_end{};
    // device walk: walkers' start lists, their counts and join indices, the
    // joined pieces (begin, length, output offset), and host staging of them
    DevBuf fe_lists, fe_meta, fe_piece;
    DevBuf fe_flag;  // jump seeding: set by a batch walk that left its piece
    DevBuf fe_stat;  // the device walk's stitch status: failed, start n
    uint64_t *fe_stat_host = nullptr; // its pinned host copy (hipHostMalloc)
    DevBuf fe_seq;   // jump seeding: batches left to the one-thread walk
    DevBuf xo_states, xo_draws; // faster-evgen on xoshiro256+: piece states and uniforms
    DevBuf xo128_states, xo128_draws, xo128_pt; // the same on xoshiro128+ (f32), and the piece matrix powers
    DevBuf xo_acc;              // accept bits of the expanded xoshiro pieces: acc1, then acc3
    DevBuf xo_lt, xo128_lt;     // the expanding lanes' skip matrices (xoshiro256+, xoshiro128+)
    uint64_t xo_acc_words = 0;  // words per array in xo_acc
    std::string xo128_pt_key;
    int xo128_pt_kb = 0;
    DevBuf adj_u;               // tp_adjoint: the stream's points, expanded
    uint32_t fe_flag_host = 0;
    uint64_t fe_qnext = 0;  // the device walk's start n: where the next chunk begins
    uint64_t fe_xo_end = 0; // faster-evgen on xoshiro256+ with skip-ahead seeding: draws the last run consumed
    uint64_t fe_host_chunks = 0; // faster-evgen: chunks of the last run walked on the host
    uint64_t fe_chunks = 0;      // faster-evgen: chunks the last run was split into
};

constexpr int KM = 32;              // intra-call offsets < 2^32 draws
constexpr int RANF_R = 2048;        // RANF rounds per expansion segment (one jump per segment)
// Batches per RANF expansion + event-loop launch: two full strides of the
// 3072-wavefront persistent grid (one wavefront per batch), so no launch ends
// on a partly idle stride; 6144 * 120000 int32 = 2.95 GB of stream.
constexpr uint64_t RANF_CHUNK_BATCHES = 6144;

// Where tp_guard stores an exception's message for tp_last_error.
inline std::string *ctx_err(void *ctx) { return ctx ? &static_cast<Ctx *>(ctx)->err : nullptr; }

inline int fail(Ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail((ctx), TP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

inline int ensure(Ctx *c, DevBuf &b, size_t bytes)
{
    if (b.n >= bytes && b.p)
        return TP_OK;
    if (b.p) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess)
        return fail(c, TP_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    b.n = bytes;
    return TP_OK;
}

// Persistent grid: as many batch workgroups as can be resident at once (CUs x
// occupancy), each striding over batches; never more than there are batches.
inline int grid_for(const Ctx *c, int source, int precision, uint64_t nb)
{
    const uint64_t g = static_cast<uint64_t>(c->cus) * static_cast<uint64_t>(c->blocks_per_cu[source][precision]);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(nb, g)));
}

// Device-side constants of a run, computed exactly as the reference does on
// the host (coupling.rs, resacc.rs:59-100, evgen.rs:66-69, spinor.rs:37-44).
tpd::Params make_params(const tp_params *p);
tpd::FParams make_fparams(const tp_params *p);

} // namespace tpe
