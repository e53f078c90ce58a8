// tp_kernels.h -- host-side launch interface of the gfx950 kernels
// (implemented in tp_kernels.hip).  Plain structs of device pointers; no
// ownership here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tp_physics.h"

namespace tpk {


// Arguments of the fused event-loop kernel.
struct FusedArgs {
    tpd::Params P;
    uint64_t nb;           // batches in this launch
    uint64_t b_first;      // global index of the first batch of this launch
    uint64_t num_events;   // events of the whole run (for the last, partial batch)
    uint32_t batch_size;   // events per full batch
    int32_t precision;     // tp_precision
    // xoshiro source: state of batch b, thread chunk start e0 is
    //   M^(12*e0) * B^b * s0  with B = M^(12*batch_size) or the 2^128 jump.
    uint64_t s0[4];
    const uint64_t *mt;    // M^(2^k), k < km, each 256x4 words
    const uint64_t *bt;    // B^(2^k), k < kb
    const uint64_t *stride;// B^(grid size)
    int32_t km, kb;
    // buffer source: RANF values (int32 in [0, 1e9)) of this launch's batches,
    // batch (b - b_first) at offset 12*batch_size*(b - b_first).
    const int32_t *draws;
    // weighted source (importance sampling): uniforms (f64) and weight 1/q(u)
    // per event; event i = batch_size*(b - b_first) + e has its uniform j at
    // u[j * u_stride + i] (one row per uniform, so a wavefront's loads are
    // contiguous) and its weight at invq[i]; every |M|^2 is multiplied by
    // its weight.
    const double *u;
    const double *invq;
    uint64_t u_stride;
    // faster-evgen source (RANF, f64): stream values from position qbase in
    // draws[], event e of batch (b - b_first) starting at stream position
    // evq[batch_size*(b - b_first) + e].
    const uint64_t *evq;
    uint64_t qbase;
    // faster-evgen from the xoshiro256+ stream (source 4): its uniforms, the
    // starts evq and qbase as above
    const double *xdraws;
    // faster-evgen: the stream's accept bits (tpd::RerollBits), valid below
    // blimit, or for pieces (bstride > 0) batch k's below (k + 1) * bstride;
    // acc1 null: re-rolls one pair at a time
    const uint32_t *acc1;
    uint64_t bbase, blimit, bstride;
    // output: one accumulator per batch, index b - b_first
    void *out;             // tp_acc*
};

// Arguments of the RANF expansion kernel (one wavefront per segment of R
// rounds).  A piece is one stream origin: the reference's master stream in
// skip-ahead mode, or one reseeded batch stream in jump mode.
struct RanfArgs {
    const int32_t *origin; // [npieces][56] round X[1..55] at [1..55]
    const int32_t *p0;     // [npieces] values of the origin round already consumed
    uint64_t d_begin, d_end; // window of draws (piece-relative) to write
    uint64_t seg_first;    // first segment index of the window
    uint64_t nsegs;        // segments per piece in this launch
    uint64_t piece_stride; // output offset between pieces
    const uint32_t *polys; // [K][55] x^(55*R*2^k) mod (x^55 + x^31 - 1)
    int32_t K;
    int32_t R;             // rounds per segment
    int32_t *out;
    // faster-evgen: accept bits of every round the launch generates (bits 1:
    // the unit-disc test in f64, 2: in f32; 0: none), as FeWalkArgs.  Piece
    // k's segment s fills the words from k * bits_piece_words + (s - seg_first)
    // * 55 R / 32 on (55 R a multiple of 32), bit j = the segment's j-th position.
    int32_t bits;
    uint64_t bits_piece_words;
    uint32_t *acc1, *acc3;
};

void launch_fused(const FusedArgs &a,
                  int source /*0 xoshiro, 1 buffer, 2 weighted, 3 faster-evgen RANF, 4 faster-evgen xoshiro (f64 only)*/,
                  int grid, hipStream_t st);
int fused_blocks_per_cu(int source, int precision); // resident workgroups per CU (occupancy)
// Per-event diagnostics: one thread per event, same device code as the loop.
// Event e draws from offset 12*e (xoshiro: a.s0 with a.mt; buffer: a.draws).
struct EventOut {
    double *u;     // [n][12]
    double *p;     // [n][3][4]
    int32_t *keep; // [n]
    double *m;     // [n][5]
};
void launch_events(const FusedArgs &a, int source, uint32_t n, const EventOut &o, hipStream_t st);
void launch_ranf_expand(const RanfArgs &a, uint64_t npieces, hipStream_t st);
// Per-batch RANF origins of jump mode: seeded_new(seed0 + 123456*b).
void launch_ranf_seed(int32_t seed0, uint64_t b_first, uint64_t nb, int32_t *origin, int32_t *p0,
                      hipStream_t st);
// In-order left fold of nb accumulators (resacc.rs:133-139) into *merged.
// In-order fold of nb accumulators; f32 = 1 adds in float (all-f32 mode).
void launch_merge(const void *per_batch, uint64_t nb, void *merged, hipStream_t st, int f32 = 0);

// ---- all-f32 mode: events, then per-batch sums in event order --------------
struct F32Args {
    tpd::FParams P;
    uint64_t nb, b_first, num_events;
    uint32_t batch_size;
    int32_t source;        // 0 xoshiro128+ in registers, 1 RANF values from draws, 2 faster-evgen on
                           // RANF draws, 3 faster-evgen on xoshiro128+ uniforms (x32draws)
    uint32_t s0[4];        // xoshiro128+ state of the run
    const uint32_t *mt;    // M^(2^k) of xoshiro128+, k < km (128 columns x 4 words)
    const uint32_t *bt;    // B^(2^k), k < kb: B = M^(12 batch_size) or the 2^64 jump
    const uint32_t *stride;// B^stride_grid: a lane's next batch of a grid-stride loop
    int32_t km, kb;
    int32_t stride_grid;
    const int32_t *draws;  // RANF values of this launch's batches, 12 per event, batch-major
    // faster-evgen (source 2, RANF): event starts evq[batch_size*(b - b_first) + e]
    // in draws from position qbase, as FusedArgs
    const uint64_t *evq;
    uint64_t qbase;
    const float *x32draws; // faster-evgen (source 3): xoshiro128+ uniforms from position qbase
    const uint32_t *acc1;  // faster-evgen: accept bits, as FusedArgs
    uint64_t bbase, blimit, bstride;
    float *ev;             // per event m[5] and keep (0/1), at batch_size*(b - b_first) + e
    void *out;             // tp_acc per batch, index b - b_first (f32 values)
};
void launch_f32_events(const F32Args &a, int grid, hipStream_t st);
void launch_f32_accumulate(const F32Args &a, int grid, hipStream_t st);

// ---- faster-evgen stream walk on the device ----------------------------------
// Event starts of the faster-evgen draws (evgen.rs:209-255, random_array's
// discarded round leftovers ranf.rs:89-92), split like the host walk: walker
// k starts at seg_origin + k*seg (walker 0 at the true start q_start), as if
// an event started there, and records its event starts until it passes the
// next segment's start by `overlap` positions.  Positions are counted from
// the start of the master state's round; buf holds positions [qbase, qlimit).
struct FeWalkArgs {
    const int32_t *buf;    // RANF values of positions qbase .. qlimit - 1 (kinds 0, 1)
    int32_t kind;          // 0: RANF, unit-disc test in f64; 1: RANF, in f32; 2: xoshiro256+; 3: xoshiro128+ (f32)
    // The walk reads accept bits, not values.  Bit i of acc1 (word i / 32,
    // bit i % 32) is 1 when the point (2u - 1) of positions bbase + i and
    // bbase + i + 1 passes the unit-disc test; acc3 likewise for bbase + i and
    // bbase + i + 3 (the three points of each event's 6 values).  buf (RANF)
    // still serves the one-thread walk of a batch the segments leave uncovered.
    const uint32_t *acc1, *acc3;
    uint64_t bbase;        // position of bit 0 of acc1 / acc3
    uint64_t qbase, qlimit;
    uint64_t q_start;
    uint64_t seg_origin, seg, overlap;
    uint32_t nwalkers;
    uint32_t cap;          // list capacity per walker (>= its longest possible walk)
    uint64_t *lists;       // [nwalkers][cap] event starts, ascending
    uint32_t *counts;      // [nwalkers]
    int32_t *join_prev;    // [nwalkers] for k >= 1: index in list k-1 of the first start >= the
    int32_t *join_next;    //   start of segment k that list k also holds, and its index there (-1: none)
};
void launch_fe_walk(const FeWalkArgs &a, hipStream_t st);
// Copies the pieces of the joined walk into out: piece k is lists[k][begin[k] ..
// begin[k] + len[k]) at out[off[k] ..].
// The chaining of the joined walk on the device: piece k of the first `need`
// starts is lists[k][begin[k] .. begin[k] + len[k]) at out[off[k] ..];
// status[0] = 1 when some walker does not meet its successor after its own
// join, or the pieces hold fewer than `need` starts; else status[1] = start
// need - 1.  status holds FE_STITCH_STATUS words (the rest: per-block sums and
// flags); at most 256 * FE_STITCH_THREADS walkers.
constexpr int FE_STITCH_THREADS = 1024;
constexpr int FE_STITCH_STATUS = 2 + 2 * FE_STITCH_THREADS;
void launch_fe_stitch(const FeWalkArgs &a, uint64_t need, uint32_t *begin, uint32_t *len, uint64_t *off,
                      uint64_t *status, hipStream_t st);
void launch_fe_scatter(const FeWalkArgs &a, const uint32_t *begin, const uint32_t *len, const uint64_t *off,
                       uint64_t *out, hipStream_t st);

// faster-evgen with jump seeding (faster-threading): every batch reseeds, so
// each batch's stream is its own piece and is walked by one thread from the
// piece's first value.  Piece k (batch b_first + k) holds the frame positions
// q of its state (q counted from the start of the state's round) at
// buf[k * piece_stride + q], piece_stride a multiple of 55, so that the
// global position G = k * piece_stride + q keeps q's place in its round
// (G % 55 = q % 55).  evq[k * batch_size + e] receives event e's start as G.
// A walk that would leave its piece sets *overflow and fills the rest of its
// batch with the piece's first position (every later read stays in bounds).
// The xoshiro pieces (kinds 2, 3) hold uniforms from position 0 of each
// batch's own stream (p0 null), and the walk reads their accept bits.
//
// Each batch is walked by S segment walkers, as the one-stream walk is
// (walker s of batch k starts at k * piece_stride + s * seg, walker 0 at the
// batch's first position, and walks overlap positions past its segment),
// joined where they meet; a batch whose walkers do not all meet before its
// events are covered is walked by one thread from its first position.
struct FeBatchWalkArgs {
    const int32_t *buf;    // RANF pieces (kinds 0, 1)
    int32_t kind;          // 0: RANF, unit-disc test in f64; 1: in f32 (all-f32 mode); 2: xoshiro256+;
                           // 3: xoshiro128+ (all-f32 mode)
    const uint32_t *acc1, *acc3; // accept bits of the pieces' positions G, as FeWalkArgs (bbase 0)
    uint64_t piece_stride;
    const int32_t *p0;     // [nb] frame position of each piece's first value (RANF)
    uint64_t nb, b_first, num_events;
    uint32_t batch_size;
    uint64_t *evq;
    uint32_t *overflow;
    uint32_t S;            // walkers per batch
    uint64_t seg, overlap; // segment length (a multiple of 55 for RANF) and overlap
    uint32_t cap;          // list capacity per walker
    uint64_t *lists;       // [nb * S][cap]
    uint32_t *counts;      // [nb * S]
    int32_t *join_prev, *join_next; // [nb * S] as FeWalkArgs, within each batch
    uint32_t *seq;         // [nb] 1: walk the batch with one thread
};
void launch_fe_walk_batches(const FeBatchWalkArgs &a, hipStream_t st);

// Pieces of the xoshiro256+ stream: piece k starts from B^(b_first + k) s0
// (B the 2^128 jump with jump seeding, each piece a batch's stream; or
// M^piece_len, the pieces tiling one stream) and holds its first piece_len
// uniforms at xbuf[k * piece_len ..], and the accept bits of those positions
// at acc1 / acc3 (the last three positions' bits from the draws that follow
// the piece: the next piece's first ones when the pieces tile one stream).
struct XoPieceArgs {
    uint64_t s0[4];
    const uint64_t *bt;    // B^(2^k), k < kb
    int32_t kb;
    const uint64_t *lt;    // M^(per_thread 2^j), j < 6: lane l's skip of l * per_thread draws
    uint64_t nb, b_first;
    uint64_t *states;      // [nb][4] scratch: each piece's first state
    uint64_t piece_len;    // uniforms expanded per piece: 64 * per_thread
    uint32_t per_thread;   // uniforms per lane, a multiple of 128
    double *xbuf;          // [nb][piece_len]
    uint32_t *acc1, *acc3; // [nb][piece_len / 32] accept bits of the positions, as FeWalkArgs
};
void launch_xo_expand(const XoPieceArgs &a, hipStream_t st); // piece states and expansion

// The same for the xoshiro128+ stream of the all-f32 mode (B the 2^64 jump
// or M^piece_len; f32 uniforms).
struct Xo128PieceArgs {
    uint32_t s0[4];
    const uint32_t *bt;    // B^(2^k), k < kb (128 columns x 4 words each)
    int32_t kb;
    const uint32_t *lt;    // M^(per_thread 2^j), j < 6
    uint64_t nb, b_first;
    uint32_t *states;      // [nb][4] scratch: each piece's first state
    uint64_t piece_len;    // 64 * per_thread
    uint32_t per_thread;   // a multiple of 128
    float *xbuf;           // [nb][piece_len]
    uint32_t *acc1, *acc3; // [nb][piece_len / 32] accept bits (the unit-disc test in f32)
};
void launch_xo128_expand(const Xo128PieceArgs &a, hipStream_t st);

} // namespace tpk

namespace tpk {

// Arguments of the adjoint kernel (tp_adjoint).
// The adjoint kernels run one 64-lane workgroup per 64 consecutive points,
// one point per lane; the f32 kernel stores the rows through LDS
// (tp_adjstore.h), the f64 kernel straight from registers.  Points of the
// xoshiro stream (u_in null) are first expanded into u_gen by a kernel whose
// threads each step through ADJ_GEN_POINTS consecutive points from one
// skip-ahead.
struct AdjArgs {
    tpd::Params P;
    double w;              // RAMBO event weight
    uint64_t n;            // points
    // uniforms: u_in (n x 12 doubles) if non-null, else xoshiro from s0 via mt,
    // expanded into u_gen (n x 12 doubles)
    const double *u_in;
    double *u_gen;
    uint64_t s0[4];
    const uint64_t *mt;
    int32_t km;
    void *F, *dF_du, *dF_dp; // [n][5], [n][5][12], [n][5][12] of double or float
};
constexpr uint32_t ADJ_GEN_POINTS = 256;

void launch_adjoint(const AdjArgs &a, int precision /*TP_PREC_F64 or TP_PREC_F32*/, hipStream_t st);

} // namespace tpk
