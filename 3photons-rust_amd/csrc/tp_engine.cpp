// tp_engine.cpp -- C ABI of the engine (include/tp.h): contexts, stream
// tables, and the orchestration of the gfx950 kernels for one batch
// (tp_simulate_batch) or one whole run (tp_run / tp_run_device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "tp.h"
#include "tp_host.h"
#include "tp_kernels.h"
#include "tp_rng_host.h"
#include "tp_ctx.h"

namespace tpe {
tpd::Params make_params(const tp_params *p)
{
    tpd::Params P;
    std::memset(&P, 0, sizeof P);
    P.e_total = p->e_total;
    P.cut_beam = p->cut_beam;
    P.cut_pp = p->cut_pp;
    P.e_min = p->e_min;
    P.cut_plane = p->cut_plane;
    P.g_a = p->g_a;
    P.g_bp = p->g_beta_p;
    P.g_bm = p->g_beta_m;
    for (int i = 0; i < 5; ++i)
        P.sc[i] = p->sigma_contribs[i];
    // Incoming momenta (evgen.rs:66-69) and their spinor pieces (spinor.rs:37-44).
    const double e = p->e_total;
    P.el_x = -e / 2.;
    P.el_e = e / 2.;
    const double pos_x = e / 2., pos_e = e / 2.;
    P.xx0 = std::sqrt(P.el_e + 0.);
    P.f0re = P.el_x / P.xx0;
    P.f0im = 0. / P.xx0;
    P.xx1 = std::sqrt(pos_e + 0.);
    P.f1re = pos_x / P.xx1;
    P.f1im = 0. / P.xx1;
    const double s01re = P.f0re * P.xx1 - P.f1re * P.xx0;
    const double s01im = P.f0im * P.xx1 - P.f1im * P.xx0;
    const double N = -(2. * 1.41421356237309504880); // -RAC8 (spinor.rs:19)
    P.ns01re = N * s01re;
    P.ns01im = N * s01im;
    P.nt01re = N * -s01re; // t = -conj(s)
    P.nt01im = N * s01im;
    // The beams run along x: f0im = f1im = 0, so s01im = 0 and ns01, nt01 are
    // real; the amplitudes scale by these products of them and the couplings.
    P.ka_s = P.ns01re * P.g_a;
    P.ka_t = P.nt01re * P.g_a;
    P.kp_s = P.ns01re * P.g_bp;
    P.kp_t = P.nt01re * P.g_bp;
    P.km_s = P.ns01re * P.g_bm;
    P.km_t = P.nt01re * P.g_bm;
    P.sort = p->sort_photons ? 1 : 0;
    return P;
}

// The same constants in f32 arithmetic for the all-f32 mode (the tp_params
// values are f32 values there, from tp_params_from_config_f32).
tpd::FParams make_fparams(const tp_params *p)
{
    tpd::FParams F;
    std::memset(&F, 0, sizeof F);
    F.e_total = (float)p->e_total;
    F.cut_beam = (float)p->cut_beam;
    F.cut_pp = (float)p->cut_pp;
    F.e_min = (float)p->e_min;
    F.cut_plane = (float)p->cut_plane;
    F.g_a = (float)p->g_a;
    F.g_bp = (float)p->g_beta_p;
    F.g_bm = (float)p->g_beta_m;
    for (int i = 0; i < 5; ++i)
        F.sc[i] = (float)p->sigma_contribs[i];
    const float e = F.e_total;
    F.el_x = -e / 2.f;
    F.el_e = e / 2.f;
    const float pos_x = e / 2.f, pos_e = e / 2.f;
    F.xx0 = std::sqrt(F.el_e + 0.f);
    F.f0re = F.el_x / F.xx0;
    F.f0im = 0.f / F.xx0;
    F.xx1 = std::sqrt(pos_e + 0.f);
    F.f1re = pos_x / F.xx1;
    F.f1im = 0.f / F.xx1;
    const float s01re = F.f0re * F.xx1 - F.f1re * F.xx0;
    const float s01im = F.f0im * F.xx1 - F.f1im * F.xx0;
    const float N = -(2.f * 1.41421356237309504880f); // -RAC8 in f32 (spinor.rs:19)
    F.ns01re = N * s01re;
    F.ns01im = N * s01im;
    F.nt01re = N * -s01re;
    F.nt01im = N * s01im;
    F.sort = p->sort_photons ? 1 : 0;
    return F;
}

} // namespace tpe

namespace {

using namespace tpe;

int bits_for(uint64_t n);

// xoshiro128+ skip-ahead tables for the all-f32 mode: M^(2^k) (k < KM128:
// a lane's draw offset inside its batch, below 12 * batch_size) and B^(2^k)
// with B = M^(12 batch_size) or the 2^64 jump.  The largest batch is one
// tp_simulate_batch call of 640 000 events: offsets below 7 680 000 < 2^23.
constexpr int KM128 = 24;
static_assert(12ull * 64 * TP_BATCH_SIZE < (1ull << KM128), "xoshiro128+ offset table too short");

int ensure_mt128(Ctx *c)
{
    if (!c->mt128.p) {
        const std::vector<tp::Gf2Mat128> &t = tp::xoshiro128_step_pow2();
        int rc = ensure(c, c->mt128, KM128 * sizeof(tp::Gf2Mat128));
        if (rc)
            return rc;
        HIPCHK(c, hipMemcpy(c->mt128.p, t.data(), KM128 * sizeof(tp::Gf2Mat128), hipMemcpyHostToDevice));
    }
    return TP_OK;
}

// B^(2^k), k < kb, of the xoshiro128+ piece matrix B (the 2^64 jump, or
// M^piece_len for pieces tiling one stream), into c->xo128_pt.
int ensure_xo128_piece_table(Ctx *c, bool jump, uint64_t piece_len, uint64_t max_piece)
{
    const int kb = std::max(1, bits_for(max_piece + 1));
    const std::string key = jump ? std::string("jump") : "piece:" + std::to_string(piece_len);
    if (c->xo128_pt_key == key && c->xo128_pt_kb >= kb)
        return TP_OK;
    tp::Gf2Mat128 B;
    if (jump)
        tp::gf2_128_jump64(B);
    else
        tp::gf2_128_pow_from_table(tp::xoshiro128_step_pow2(), piece_len, B);
    const std::vector<tp::Gf2Mat128> bt = tp::gf2_128_pow2_table(B, kb);
    int rc = ensure(c, c->xo128_pt, bt.size() * sizeof(tp::Gf2Mat128));
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->xo128_pt.p, bt.data(), bt.size() * sizeof(tp::Gf2Mat128), hipMemcpyHostToDevice));
    c->xo128_pt_key = key;
    c->xo128_pt_kb = kb;
    return TP_OK;
}

int ensure_tables128(Ctx *c, int seeding, uint32_t batch_size, uint64_t nb, int grid, DevBuf **bt_out, int *kb_out)
{
    int rc0 = ensure_mt128(c);
    if (rc0)
        return rc0;
    const int kb = std::max(1, bits_for(nb));
    const std::string key = std::to_string(seeding) + ":" + std::to_string(batch_size);
    if (c->bt128_key != key || c->bt128_kb < kb) {
        tp::Gf2Mat128 B;
        if (seeding == TP_SEED_JUMP)
            tp::gf2_128_jump64(B);
        else
            tp::gf2_128_pow_from_table(tp::xoshiro128_step_pow2(), 12ull * batch_size, B);
        const std::vector<tp::Gf2Mat128> bt = tp::gf2_128_pow2_table(B, kb);
        int rc = ensure(c, c->bt128, bt.size() * sizeof(tp::Gf2Mat128));
        if (rc)
            return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->bt128.p, bt.data(), bt.size() * sizeof(tp::Gf2Mat128), hipMemcpyHostToDevice));
        c->bt128_key = key;
        c->bt128_kb = kb;
    }
    const std::string skey = key + ":" + std::to_string(grid);
    if (c->st128_key != skey) {
        tp::Gf2Mat128 B, S;
        if (seeding == TP_SEED_JUMP)
            tp::gf2_128_jump64(B);
        else
            tp::gf2_128_pow_from_table(tp::xoshiro128_step_pow2(), 12ull * batch_size, B);
        tp::gf2_128_pow_from_table(tp::gf2_128_pow2_table(B, std::max(1, bits_for((uint64_t)grid))), (uint64_t)grid, S);
        int rc = ensure(c, c->st128, sizeof S);
        if (rc)
            return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->st128.p, &S, sizeof S, hipMemcpyHostToDevice));
        c->st128_key = skey;
    }
    *bt_out = &c->bt128;
    *kb_out = c->bt128_kb;
    return TP_OK;
}

// 24576 * 10000 events * 24 B = 5.9 GB of per-event values.  The grid is 24
// workgroups per CU (6 waves per SIMD, the occupancy the kernel's register
// budget allows), so each lane steps through 4 batches of a full chunk by the
// grid-stride matrix instead of a full skip-ahead per batch.
constexpr uint64_t F32_CHUNK_BATCHES = 24576;

int f32_grid(const Ctx *c, uint64_t nb) { return (int)std::max<uint64_t>(1, std::min<uint64_t>(nb, 24ull * c->cus)); }

// All-f32 run with the xoshiro128+ stream.
int enqueue_xo128_f32(Ctx *c, const tp_params *params, int seeding, const tp_xoshiro128_state *x,
                      uint64_t num_events, uint32_t batch_size, void *per_batch_dev)
{
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    if (12ull * batch_size >= (1ull << KM128))
        return fail(c, TP_ERR_ARG, "all-f32 batch exceeds the xoshiro128+ offset range");
    DevBuf *bt = nullptr;
    int kb = 0;
    const int sgrid = f32_grid(c, F32_CHUNK_BATCHES);
    int rc = ensure_tables128(c, seeding, batch_size, nb, sgrid, &bt, &kb);
    if (rc)
        return rc;
    const uint64_t chunk = std::min<uint64_t>(nb, F32_CHUNK_BATCHES);
    rc = ensure(c, c->f32ev, chunk * batch_size * 6 * sizeof(float));
    if (rc)
        return rc;
    tpk::F32Args a;
    std::memset(&a, 0, sizeof a);
    a.P = make_fparams(params);
    a.num_events = num_events;
    a.batch_size = batch_size;
    a.source = 0;
    std::memcpy(a.s0, x->s, sizeof a.s0);
    a.mt = static_cast<const uint32_t *>(c->mt128.p);
    a.km = KM128;
    a.bt = static_cast<const uint32_t *>(bt->p);
    a.kb = kb;
    a.stride = static_cast<const uint32_t *>(c->st128.p);
    a.stride_grid = sgrid;
    a.ev = static_cast<float *>(c->f32ev.p);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        a.nb = std::min(nb, b0 + chunk) - b0;
        a.b_first = b0;
        a.out = static_cast<tp_acc *>(per_batch_dev) + b0;
        tpk::launch_f32_events(a, f32_grid(c, a.nb), c->stream);
        HIPCHK(c, hipGetLastError());
        tpk::launch_f32_accumulate(a, f32_grid(c, a.nb), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    return TP_OK;
}


int upload_mats(Ctx *c, DevBuf &b, const std::vector<tp::Gf2Mat> &mats)
{
    size_t bytes = std::max<size_t>(1, mats.size()) * sizeof(tp::Gf2Mat);
    int rc = ensure(c, b, bytes);
    if (rc)
        return rc;
    if (!mats.empty())
        HIPCHK(c, hipMemcpy(b.p, mats.data(), mats.size() * sizeof(tp::Gf2Mat), hipMemcpyHostToDevice));
    return TP_OK;
}

int ensure_mt(Ctx *c)
{
    if (c->km)
        return TP_OK;
    const std::vector<tp::Gf2Mat> &t = tp::xoshiro_step_pow2();
    std::vector<tp::Gf2Mat> v(t.begin(), t.begin() + KM);
    int rc = upload_mats(c, c->mt, v);
    if (rc)
        return rc;
    c->km = KM;
    return TP_OK;
}

int bits_for(uint64_t n)
{
    int k = 0;
    while (k < 64 && (n >> k))
        ++k;
    return k;
}

// Batch-advance matrix B (skip-ahead: M^(12*batch_size); jump: the 2^128
// jump) as pow2 table up to nb, plus B^grid.
int ensure_tables(Ctx *c, int seeding, uint32_t batch_size, uint64_t nb, int grid, Tables **out)
{
    int kb = std::max(1, bits_for(nb));
    std::string key = std::to_string(seeding) + ":" + std::to_string(batch_size) + ":" + std::to_string(grid);
    auto it = c->tables.find(key);
    if (it != c->tables.end() && it->second->kb >= kb) {
        *out = it->second.get();
        return TP_OK;
    }
    tp::Gf2Mat B;
    if (seeding == TP_SEED_JUMP)
        tp::gf2_jump128(B);
    else
        tp::gf2_pow_from_table(tp::xoshiro_step_pow2(), 12ull * batch_size, B);
    std::vector<tp::Gf2Mat> bt = tp::gf2_pow2_table(B, kb);
    tp::Gf2Mat S;
    tp::gf2_pow_from_table(bt, static_cast<uint64_t>(grid), S);
    if (bits_for(static_cast<uint64_t>(grid)) > kb) {
        std::vector<tp::Gf2Mat> bt2 = tp::gf2_pow2_table(B, bits_for(static_cast<uint64_t>(grid)));
        tp::gf2_pow_from_table(bt2, static_cast<uint64_t>(grid), S);
    }
    std::unique_ptr<Tables> t(new Tables);
    int rc = upload_mats(c, t->bt, bt);
    if (rc)
        return rc;
    rc = upload_mats(c, t->stride, std::vector<tp::Gf2Mat>(1, S));
    if (rc)
        return rc;
    t->kb = kb;
    *out = t.get();
    c->tables[key] = std::move(t);
    return TP_OK;
}

// x^(55*R*2^k) mod (x^55 + x^31 - 1) for k < K, uploaded to c->polys.
int ensure_polys(Ctx *c, int R, int K)
{
    std::string key = std::to_string(R) + ":" + std::to_string(K);
    auto it = c->ranf_polys.find(key);
    if (it == c->ranf_polys.end()) {
        std::vector<uint32_t> flat;
        tp::RanfPoly q = tp::ranf_poly_xpow(55ull * static_cast<uint64_t>(R));
        for (int k = 0; k < K; ++k) {
            flat.insert(flat.end(), q.begin(), q.end());
            q = tp::ranf_poly_mul(q, q);
        }
        it = c->ranf_polys.emplace(key, std::move(flat)).first;
    }
    int rc = ensure(c, c->polys, it->second.size() * sizeof(uint32_t));
    if (rc)
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->polys.p, it->second.data(), it->second.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->stream));
    return TP_OK;
}

int check_common(Ctx *c, const tp_params *params, int rng_kind, const void *state)
{
    if (!params || !state)
        return fail(c, TP_ERR_ARG, "null params or rng state");
    if (rng_kind != TP_RNG_RANF && rng_kind != TP_RNG_XOSHIRO && rng_kind != TP_RNG_XOSHIRO128)
        return fail(c, TP_ERR_ARG, "unknown rng kind " + std::to_string(rng_kind));
    if (params->precision != TP_PREC_F64 && params->precision != TP_PREC_MIXED && params->precision != TP_PREC_F32)
        return fail(c, TP_ERR_UNSUPPORTED, "precision " + std::to_string(params->precision) + " not available");
    if ((params->precision == TP_PREC_F32) != (rng_kind == TP_RNG_XOSHIRO128) && rng_kind != TP_RNG_RANF)
        return fail(c, TP_ERR_ARG, "xoshiro128+ is the stream of the all-f32 mode, xoshiro256+ of the others");
    if (params->faster_evgen && params->precision == TP_PREC_MIXED)
        return fail(c, TP_ERR_UNSUPPORTED, "faster-evgen is available in f64 and f32, not in mixed precision");
    if (!(params->e_total > 0.) || !std::isfinite(params->e_total))
        return fail(c, TP_ERR_ARG, "e_total must be positive and finite");
    return TP_OK;
}

// RANF with `faster-evgen`: the event starts of the reference's
// simulate_event_batch walk, then each chunk of batches is simulated from
// them with its stream expanded into HBM.  Runs of FE_DEVICE_WALK_MIN events
// or more walk the stream on the device (fe_walk_kernel: segment walkers
// joined where they meet); shorter runs, and any chunk whose device walkers
// do not all join, take the host walk.
constexpr uint64_t FE_DEVICE_WALK_MIN = 1u << 20;
constexpr uint64_t FE_WALK_BATCHES = 12288; // batches per chunk: four strides of the event-loop grid
constexpr double FE_POSITIONS_PER_EVENT = 18.5; // stream positions allowed for per event (mean 17.9)
constexpr uint64_t FE_EVENTS_PER_WALKER = 256, FE_MAX_WALKERS = 1u << 18;
static_assert(FE_MAX_WALKERS <= 256ull * tpk::FE_STITCH_THREADS, "the device stitch's block sums fit one workgroup");
// Positions a walker goes past its segment's end: 24 rounds, ~74 events; a
// walker meets the true walk within 49 events in the runs checked.
constexpr uint64_t FE_OVERLAP = 24 * 55;
// The xoshiro256+ stream has no rounds whose discards bring walkers together:
// a walker meets the true walk only where both happen to start an event, after
// a median of 22 events but a long tail (at most 832 over 176 447 walkers of a
// simulated 1.5e9-draw stream, 12 of them past 512).  Its walkers therefore
// take 1024 events each and go 1024 events (17 500 positions) past their
// segment; 16.64 positions per event on average, 17.0 allowed for.
constexpr double FE_XO_POSITIONS_PER_EVENT = 17.0;
constexpr uint64_t FE_XO_EVENTS_PER_WALKER = 1024, FE_XO_OVERLAP = 17500;
// xoshiro streams are expanded in pieces of FE_XO_PIECE uniforms, one
// wavefront per piece and FE_XO_PER_THREAD consecutive draws per lane from one
// skip-ahead (a multiple of 128: the lanes' accept-bit words go out 4 at a
// time).  With jump seeding a piece is a batch's stream; with skip-ahead
// seeding the pieces tile the run's one stream (piece p starts FE_XO_PIECE p
// draws after the run's state).
constexpr uint32_t FE_XO_PER_THREAD = 2688;
constexpr uint64_t FE_XO_PIECE = 64 * FE_XO_PER_THREAD;
constexpr uint64_t FE_XO_WORDS = FE_XO_PIECE / 32; // accept-bit words per piece and test
static_assert(FE_XO_PER_THREAD % 128 == 0, "whole groups of 4 bit words per lane");
// Accept-bit words per RANF expansion segment (each segment's bits start on
// a word of their own).
constexpr uint64_t RANF_SEG_WORDS = 55ull * RANF_R / 32;
static_assert(55ull * RANF_R % 32 == 0, "RANF segments of whole bit words");

namespace {

// Accept-bit arrays of `need` words each in c->xo_acc: acc1, then acc3 at
// c->xo_acc_words words (sized for the largest need seen, so that the split
// stays put while the buffer is reused).
int ensure_acc_words(Ctx *c, uint64_t need)
{
    const uint64_t words = std::max(need, c->xo_acc_words);
    int rc = ensure(c, c->xo_acc, 2 * words * sizeof(uint32_t));
    if (rc)
        return rc;
    c->xo_acc_words = words;
    return TP_OK;
}

// The same for np xoshiro pieces.
int ensure_acc(Ctx *c, uint64_t np) { return ensure_acc_words(c, np * FE_XO_WORDS); }

// RANF values of stream positions [q_first, q_last) into c->draws (position
// q_first at index 0), from the master origin uploaded by the caller, and
// their accept bits (the unit-disc test of stream kind 0 or 1) into c->xo_acc
// from *bbase, the first position of the first expansion segment, on.
int fe_expand(Ctx *c, uint64_t p0_master, int K, uint64_t q_first, uint64_t q_last, int kind, uint64_t *bbase)
{
    const uint64_t seg_first = q_first / 55 / RANF_R, nsegs = (q_last - 1) / 55 / RANF_R - seg_first + 1;
    int rc = ensure(c, c->draws, (q_last - q_first) * sizeof(int32_t));
    if (!rc)
        rc = ensure_acc_words(c, nsegs * RANF_SEG_WORDS);
    if (rc)
        return rc;
    tpk::RanfArgs r;
    std::memset(&r, 0, sizeof r);
    r.origin = static_cast<const int32_t *>(c->origin.p);
    r.p0 = static_cast<const int32_t *>(c->p0.p);
    r.polys = static_cast<const uint32_t *>(c->polys.p);
    r.K = K;
    r.R = RANF_R;
    r.out = static_cast<int32_t *>(c->draws.p);
    r.d_begin = q_first - p0_master;
    r.d_end = q_last - p0_master;
    r.seg_first = seg_first;
    r.nsegs = nsegs;
    r.piece_stride = 0;
    r.bits = kind == 1 ? 2 : 1;
    r.acc1 = static_cast<uint32_t *>(c->xo_acc.p);
    r.acc3 = r.acc1 + c->xo_acc_words;
    tpk::launch_ranf_expand(r, 1, c->stream);
    HIPCHK(c, hipGetLastError());
    *bbase = seg_first * 55 * RANF_R;
    return TP_OK;
}

// The stream a faster-evgen walk runs on.
struct FeStream {
    int kind = 0;                 // tpk::FeWalkArgs kind: 0 RANF f64, 1 RANF f32, 2 xoshiro256+
    uint64_t p0_master = 0;       // RANF: frame position of the master origin's next value
    int K = 0;                    // RANF: segment-jump depth of c->polys
    const uint64_t *s0 = nullptr; // xoshiro256+: the run's state (position 0)
    const uint32_t *s0_128 = nullptr; // xoshiro128+ (kind 3): the run's state (position 0)
};


// M^(FE_XO_PER_THREAD 2^j), j < 6, of both xoshiro generators: the expanding
// lanes' skips (lane l skips l * FE_XO_PER_THREAD draws into its piece).
int ensure_lane_tables(Ctx *c)
{
    if (!c->xo_lt.p) {
        tp::Gf2Mat m;
        tp::gf2_pow_from_table(tp::xoshiro_step_pow2(), FE_XO_PER_THREAD, m);
        int rc = upload_mats(c, c->xo_lt, tp::gf2_pow2_table(m, 6));
        if (rc)
            return rc;
    }
    if (!c->xo128_lt.p) {
        tp::Gf2Mat128 m;
        tp::gf2_128_pow_from_table(tp::xoshiro128_step_pow2(), FE_XO_PER_THREAD, m);
        const std::vector<tp::Gf2Mat128> t = tp::gf2_128_pow2_table(m, 6);
        int rc = ensure(c, c->xo128_lt, t.size() * sizeof(tp::Gf2Mat128));
        if (rc)
            return rc;
        HIPCHK(c, hipMemcpy(c->xo128_lt.p, t.data(), t.size() * sizeof(tp::Gf2Mat128), hipMemcpyHostToDevice));
    }
    return TP_OK;
}

// Expands xoshiro128+ pieces b_first .. b_first + np - 1 of piece_len f32
// uniforms into c->xo128_draws (piece k from B^(b_first + k) s0, B as
// ensure_xo128_piece_table made it).
int xo128_expand(Ctx *c, const uint32_t s0[4], uint64_t b_first, uint64_t np)
{
    int rc = ensure_lane_tables(c);
    if (!rc)
        rc = ensure(c, c->xo128_states, np * 4 * sizeof(uint32_t));
    if (!rc)
        rc = ensure(c, c->xo128_draws, np * FE_XO_PIECE * sizeof(float));
    if (!rc)
        rc = ensure_acc(c, np);
    if (rc)
        return rc;
    tpk::Xo128PieceArgs p;
    std::memset(&p, 0, sizeof p);
    std::memcpy(p.s0, s0, sizeof p.s0);
    p.bt = static_cast<const uint32_t *>(c->xo128_pt.p);
    p.kb = c->xo128_pt_kb;
    p.lt = static_cast<const uint32_t *>(c->xo128_lt.p);
    p.nb = np;
    p.b_first = b_first;
    p.states = static_cast<uint32_t *>(c->xo128_states.p);
    p.piece_len = FE_XO_PIECE;
    p.per_thread = FE_XO_PER_THREAD;
    p.xbuf = static_cast<float *>(c->xo128_draws.p);
    p.acc1 = static_cast<uint32_t *>(c->xo_acc.p);
    p.acc3 = p.acc1 + c->xo_acc_words;
    static_assert(FE_XO_PIECE < (1ull << KM128), "in-piece offsets within the step table");
    tpk::launch_xo128_expand(p, c->stream);
    HIPCHK(c, hipGetLastError());
    return TP_OK;
}

// (M^FE_XO_PIECE)^(2^k), k < kb, on the device: the states of the pieces.
int ensure_xo_piece_table(Ctx *c, uint64_t max_piece, Tables **out)
{
    const int kb = std::max(1, bits_for(max_piece + 1));
    const std::string key = "xo-piece:" + std::to_string(FE_XO_PIECE);
    auto it = c->tables.find(key);
    if (it != c->tables.end() && it->second->kb >= kb) {
        *out = it->second.get();
        return TP_OK;
    }
    tp::Gf2Mat B;
    tp::gf2_pow_from_table(tp::xoshiro_step_pow2(), FE_XO_PIECE, B);
    std::unique_ptr<Tables> t(new Tables);
    int rc = upload_mats(c, t->bt, tp::gf2_pow2_table(B, kb));
    if (rc)
        return rc;
    t->kb = kb;
    *out = t.get();
    c->tables[key] = std::move(t);
    return TP_OK;
}

// The stream of positions [q_first, q_last) into HBM: RANF values into
// c->draws from q_first on, or xoshiro uniforms into c->xo_draws /
// c->xo128_draws from the first piece boundary at or before q_first, with
// their accept bits in c->xo_acc.  *qbase is the position held at index 0,
// *bbase the position of bit 0.
int fe_stream_expand(Ctx *c, const FeStream &fs, uint64_t q_first, uint64_t q_last, uint64_t *qbase, uint64_t *bbase)
{
    if (fs.kind < 2) {
        *qbase = q_first;
        return fe_expand(c, fs.p0_master, fs.K, q_first, q_last, fs.kind, bbase);
    }
    const uint64_t p_first = q_first / FE_XO_PIECE, p_end = (q_last + FE_XO_PIECE - 1) / FE_XO_PIECE;
    const uint64_t np = p_end - p_first;
    if (fs.kind == 3) {
        int rc = ensure_xo128_piece_table(c, false, FE_XO_PIECE, p_end);
        if (!rc)
            rc = xo128_expand(c, fs.s0_128, p_first, np);
        if (rc)
            return rc;
        *qbase = *bbase = p_first * FE_XO_PIECE;
        return TP_OK;
    }
    Tables *t = nullptr;
    int rc = ensure_xo_piece_table(c, p_end, &t);
    if (!rc)
        rc = ensure_lane_tables(c);
    if (!rc)
        rc = ensure(c, c->xo_states, np * 4 * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->xo_draws, np * FE_XO_PIECE * sizeof(double));
    if (!rc)
        rc = ensure_acc(c, np);
    if (rc)
        return rc;
    tpk::XoPieceArgs p;
    std::memset(&p, 0, sizeof p);
    std::memcpy(p.s0, fs.s0, sizeof p.s0);
    p.bt = static_cast<const uint64_t *>(t->bt.p);
    p.kb = t->kb;
    p.lt = static_cast<const uint64_t *>(c->xo_lt.p);
    p.nb = np;
    p.b_first = p_first;
    p.states = static_cast<uint64_t *>(c->xo_states.p);
    p.piece_len = FE_XO_PIECE;
    p.per_thread = FE_XO_PER_THREAD;
    p.xbuf = static_cast<double *>(c->xo_draws.p);
    p.acc1 = static_cast<uint32_t *>(c->xo_acc.p);
    p.acc3 = p.acc1 + c->xo_acc_words;
    tpk::launch_xo_expand(p, c->stream);
    HIPCHK(c, hipGetLastError());
    *qbase = *bbase = p_first * FE_XO_PIECE;
    return TP_OK;
}

// Device walk of n + 1 event starts from the true start q_start into c->evq,
// with the stream of the walked range expanded into HBM (c->draws or
// c->xo_draws, position *qbase at index 0) and its accept bits (from
// position *bbase, valid below *blimit).  *ok is false when some walker did
// not meet its predecessor or the joined walk is too short; the caller then
// walks on the host.
int fe_device_walk(Ctx *c, const FeStream &fs, uint64_t q_start, uint64_t n, bool *ok, uint64_t *qbase,
                   uint64_t *bbase, uint64_t *blimit)
{
    *ok = false;
    const bool xo = fs.kind >= 2; // the xoshiro streams: no rounds
    tpk::FeWalkArgs w;
    std::memset(&w, 0, sizeof w);
    w.kind = fs.kind;
    w.q_start = q_start;
    if (xo) { // no rounds: segments start anywhere
        const uint64_t span = (uint64_t)std::ceil((double)(n + 1) * FE_XO_POSITIONS_PER_EVENT) + 128;
        w.nwalkers = (uint32_t)std::min<uint64_t>(FE_MAX_WALKERS, std::max<uint64_t>(1, n / FE_XO_EVENTS_PER_WALKER));
        w.seg_origin = q_start;
        w.seg = span / w.nwalkers + 1;
        w.overlap = FE_XO_OVERLAP;
    } else {
        const uint64_t span = (uint64_t)std::ceil((double)(n + 1) * FE_POSITIONS_PER_EVENT) + 55;
        w.nwalkers = (uint32_t)std::min<uint64_t>(FE_MAX_WALKERS, std::max<uint64_t>(64, n / FE_EVENTS_PER_WALKER));
        w.seg_origin = q_start / 55 * 55;
        w.seg = ((q_start - w.seg_origin + span) / w.nwalkers / 55 + 1) * 55;
        w.overlap = FE_OVERLAP;
    }
    const uint32_t T = w.nwalkers;
    // every event takes at least 15 positions: no walk holds more starts
    w.cap = (uint32_t)((w.seg + w.overlap + 63) / 15 + 2);
    w.qlimit = w.seg_origin + (uint64_t)T * w.seg + w.overlap + 128;
    int rc = fe_stream_expand(c, fs, q_start, w.qlimit, &w.qbase, &w.bbase);
    if (!rc)
        rc = ensure(c, c->fe_lists, (size_t)T * w.cap * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_meta, (size_t)T * 3 * sizeof(uint32_t));
    if (!rc)
        rc = ensure(c, c->fe_piece, (size_t)T * (2 * sizeof(uint32_t) + sizeof(uint64_t)));
    if (!rc)
        rc = ensure(c, c->evq, (n + 1) * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_stat, tpk::FE_STITCH_STATUS * sizeof(uint64_t));
    if (!rc && !c->fe_stat_host && hipHostMalloc(reinterpret_cast<void **>(&c->fe_stat_host), 2 * sizeof(uint64_t)) != hipSuccess) {
        c->fe_stat_host = nullptr;
        rc = fail(c, TP_ERR_NOMEM, "hipHostMalloc of the walk status failed");
    }
    if (rc)
        return rc;
    w.buf = static_cast<const int32_t *>(c->draws.p);
    w.acc1 = static_cast<const uint32_t *>(c->xo_acc.p);
    w.acc3 = w.acc1 + c->xo_acc_words;
    w.lists = static_cast<uint64_t *>(c->fe_lists.p);
    w.counts = static_cast<uint32_t *>(c->fe_meta.p);
    w.join_prev = reinterpret_cast<int32_t *>(w.counts + T);
    w.join_next = w.join_prev + T;
    tpk::launch_fe_walk(w, c->stream);
    HIPCHK(c, hipGetLastError());
    // The true walk, chained on the device: walker 0's list from its start,
    // walker k's from the index where walker k-1 met it, which must come after
    // walker k-1 itself joined; the first n + 1 starts go to evq.  The host
    // reads the status (and start n, where the next chunk begins) while the
    // copies run.
    uint32_t *dbeg = static_cast<uint32_t *>(c->fe_piece.p), *dlen = dbeg + T;
    uint64_t *doff = reinterpret_cast<uint64_t *>(dlen + T);
    uint64_t *dstat = static_cast<uint64_t *>(c->fe_stat.p);
    tpk::launch_fe_stitch(w, n + 1, dbeg, dlen, doff, dstat, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->fe_stat_host, dstat, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    tpk::launch_fe_scatter(w, dbeg, dlen, doff, static_cast<uint64_t *>(c->evq.p), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(c->ev[4]));
    if (c->fe_stat_host[0]) // a failed walk's scatter copies only pieces that lie in their lists
        return TP_OK;
    c->fe_qnext = c->fe_stat_host[1];
    *qbase = w.qbase;
    *bbase = w.bbase;
    *blimit = w.qlimit;
    *ok = true;
    return TP_OK;
}

} // namespace

// Chunks of FE_WALK_BATCHES batches on one faster-evgen stream, from the event
// start q_start: each chunk's event starts (device walk for runs of
// FE_DEVICE_WALK_MIN events or more, else -- or when the device walkers do not
// all join -- host_walk), then its events: fused_kernel<3> (RANF) or <4>
// (xoshiro256+), or in f32 (fa non-null) f32_events_kernel<2> and the
// in-order sums.  *q_end receives the position after the last event.
using FeHostWalk = std::function<void(uint64_t q, uint64_t n, uint64_t *starts, uint64_t *q_next)>;
int enqueue_fe_chunks(Ctx *c, const FeStream &fs, const FeHostWalk &host_walk, tpk::FusedArgs &a, tpk::F32Args *fa,
                      uint64_t q_start, uint64_t num_events, uint32_t batch_size, void *per_batch_dev, uint64_t *q_end)
{
    const bool device_walk = num_events >= FE_DEVICE_WALK_MIN;
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    const uint64_t chunk = std::min<uint64_t>(nb, FE_WALK_BATCHES);
    c->fe_host_chunks = 0;
    if (fa) {
        int rc = ensure(c, c->f32ev, chunk * batch_size * 6 * sizeof(float));
        if (rc)
            return rc;
        fa->ev = static_cast<float *>(c->f32ev.p);
    }
    const int source = fs.kind == 2 ? 4 : 3;
    bool first = true;
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        ++c->fe_chunks;
        const uint64_t b1 = std::min(nb, b0 + chunk);
        const uint64_t e0 = b0 * batch_size, e1 = std::min<uint64_t>(b1 * batch_size, num_events);
        const uint64_t n = e1 - e0;
        if (first)
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        bool ok = false;
        uint64_t q_next = 0, qbase = 0, bbase = 0, blimit = 0;
        if (device_walk) {
            int rc = fe_device_walk(c, fs, q_start, n, &ok, &qbase, &bbase, &blimit);
            if (rc)
                return rc;
            q_next = c->fe_qnext;
        }
        if (!ok) { // host walk of this chunk from q_start
            ++c->fe_host_chunks;
            c->fe_starts.resize(n);
            host_walk(q_start, n, c->fe_starts.data(), &q_next);
            int rc = ensure(c, c->evq, n * sizeof(uint64_t));
            if (!rc)
                rc = fe_stream_expand(c, fs, q_start, q_next, &qbase, &bbase);
            if (rc)
                return rc;
            blimit = q_next;
            HIPCHK(c, hipMemcpyAsync(c->evq.p, c->fe_starts.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice,
                                     c->stream));
        }
        if (first) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            first = false;
        }
        if (fa) {
            fa->nb = b1 - b0;
            fa->b_first = b0;
            fa->draws = static_cast<const int32_t *>(c->draws.p);
            fa->x32draws = static_cast<const float *>(c->xo128_draws.p);
            fa->qbase = qbase;
            fa->acc1 = static_cast<const uint32_t *>(c->xo_acc.p);
            fa->bbase = bbase;
            fa->blimit = blimit;
            fa->bstride = 0;
            fa->evq = static_cast<const uint64_t *>(c->evq.p);
            fa->out = static_cast<tp_acc *>(per_batch_dev) + b0;
            tpk::launch_f32_events(*fa, f32_grid(c, fa->nb), c->stream);
            HIPCHK(c, hipGetLastError());
            tpk::launch_f32_accumulate(*fa, f32_grid(c, fa->nb), c->stream);
            HIPCHK(c, hipGetLastError());
        } else {
            a.nb = b1 - b0;
            a.b_first = b0;
            a.draws = static_cast<const int32_t *>(c->draws.p);
            a.xdraws = static_cast<const double *>(c->xo_draws.p);
            a.qbase = qbase;
            a.acc1 = static_cast<const uint32_t *>(c->xo_acc.p);
            a.bbase = bbase;
            a.blimit = blimit;
            a.bstride = 0;
            a.evq = static_cast<const uint64_t *>(c->evq.p);
            a.out = static_cast<tp_acc *>(per_batch_dev) + b0;
            tpk::launch_fused(a, source, grid_for(c, source, 0, a.nb), c->stream);
            HIPCHK(c, hipGetLastError());
        }
        if (!ok) // fe_starts is staged for the copy above
            HIPCHK(c, hipStreamSynchronize(c->stream));
        q_start = q_next;
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    *q_end = q_start;
    return TP_OK;
}

int enqueue_ranf_faster_evgen(Ctx *c, tpk::FusedArgs &a, tpk::F32Args *fa, const tp_ranf_state *rs,
                              uint64_t num_events, uint32_t batch_size, void *per_batch_dev)
{
    HIPCHK(c, hipStreamSynchronize(c->stream)); // the staging vectors below are reused
    tp::RanfState g0;
    std::memcpy(g0.numbers, rs->numbers, sizeof g0.numbers);
    g0.index = rs->index;
    g0.seed = rs->seed;
    FeStream fs;
    fs.kind = fa ? 1 : 0;
    fs.p0_master = 55u - rs->index; // position of the next value
    // segment jumps cover positions up to 30 per event (the mean is 17.9)
    fs.K = std::max(1, bits_for((fs.p0_master + 30 * (num_events + 1) + 4 * FE_OVERLAP) / 55 / RANF_R + 2));
    int rc = ensure_polys(c, RANF_R, fs.K);
    if (!rc)
        rc = ensure(c, c->origin, 56 * sizeof(int32_t));
    if (!rc)
        rc = ensure(c, c->p0, sizeof(int32_t));
    if (rc)
        return rc;
    std::memcpy(c->origin0, rs->numbers, sizeof c->origin0);
    c->p00 = static_cast<int32_t>(fs.p0_master);
    HIPCHK(c, hipMemcpyAsync(c->origin.p, c->origin0, sizeof c->origin0, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->p0.p, &c->p00, sizeof c->p00, hipMemcpyHostToDevice, c->stream));
    const bool f32 = fa != nullptr;
    const FeHostWalk walk = [&](uint64_t q, uint64_t n, uint64_t *starts, uint64_t *q_next) {
        tp::ranf_walk_faster_evgen_from(g0, q, n, starts, q_next, 16, f32);
    };
    uint64_t q_end = 0;
    rc = enqueue_fe_chunks(c, fs, walk, a, fa, fs.p0_master, num_events, batch_size, per_batch_dev, &q_end);
    if (rc)
        return rc;
    c->fe_end = g0;
    tp::ranf_advance(c->fe_end, q_end - fs.p0_master);
    return TP_OK;
}

// xoshiro256+ with `faster-evgen` and skip-ahead seeding (standard-random,
// faster-evgen): one stream from the run's state, random_array::<N> taking N
// consecutive draws (standard.rs:33-35), walked like the RANF stream above
// with walkers of FE_XO_EVENTS_PER_WALKER events.  c->fe_xo_end receives the
// number of draws the run consumed.
int enqueue_xo_faster_evgen(Ctx *c, tpk::FusedArgs &a, const tp_xoshiro_state *x, uint64_t num_events,
                            uint32_t batch_size, void *per_batch_dev)
{
    HIPCHK(c, hipStreamSynchronize(c->stream)); // the staging vectors below are reused
    uint64_t s0[4];
    std::memcpy(s0, x->s, sizeof s0);
    FeStream fs;
    fs.kind = 2;
    fs.s0 = s0;
    const FeHostWalk walk = [&](uint64_t q, uint64_t n, uint64_t *starts, uint64_t *q_next) {
        tp::xoshiro_walk_faster_evgen(s0, q, n, starts, q_next);
    };
    uint64_t q_end = 0;
    int rc = enqueue_fe_chunks(c, fs, walk, a, nullptr, 0, num_events, batch_size, per_batch_dev, &q_end);
    if (rc)
        return rc;
    c->fe_xo_end = q_end;
    return TP_OK;
}

// RANF with `faster-evgen` and jump seeding (faster-threading): batch b
// reseeds (seeded_new(seed + 123456 b), multi_threading.rs:64-69 and
// ranf.rs:135-140; batch 0 runs on the caller's state), so its stream is a
// piece of its own: each chunk of batches is expanded piece by piece, walked
// one thread per batch, and simulated from those starts.  A batch consumes
// about 179 000 positions (17.9 per event, at most a few hundred more than
// that in any batch of 10 000 events); a piece holds 220 000.
constexpr uint64_t FE_JUMP_PIECE = 55 * 4000;
static_assert(FE_JUMP_PIECE % 55 == 0, "pieces keep every position's place in its round");
static_assert(FE_JUMP_PIECE % 32 == 0, "pieces of whole accept-bit words");

// Segment walkers for the batches of a chunk (jump seeding): S walkers per
// batch of the stream kind's length and overlap, their lists and join
// indices, and the per-batch flags of the one-thread walk.
int fe_batch_walk_setup(Ctx *c, tpk::FeBatchWalkArgs &w, int kind, uint64_t max_batches, uint32_t batch_size)
{
    const bool xo = kind >= 2; // the xoshiro streams: no rounds
    const uint64_t per_walker = xo ? FE_XO_EVENTS_PER_WALKER : FE_EVENTS_PER_WALKER;
    w.kind = kind;
    w.S = (uint32_t)std::max<uint64_t>(1, (batch_size + per_walker - 1) / per_walker);
    const uint64_t span =
        (uint64_t)std::ceil((double)batch_size * (xo ? FE_XO_POSITIONS_PER_EVENT : FE_POSITIONS_PER_EVENT)) + 128;
    w.seg = xo ? span / w.S + 1 : (span / w.S / 55 + 1) * 55; // > 55: past every piece's first position
    w.overlap = xo ? FE_XO_OVERLAP : FE_OVERLAP;
    // every event takes at least 15 positions: no walk holds more starts
    w.cap = (uint32_t)((w.seg + w.overlap + 63) / 15 + 2);
    const uint64_t walkers = max_batches * w.S;
    int rc = ensure(c, c->fe_lists, walkers * w.cap * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_meta, walkers * 3 * sizeof(uint32_t));
    if (!rc)
        rc = ensure(c, c->fe_seq, max_batches * sizeof(uint32_t));
    if (rc)
        return rc;
    w.lists = static_cast<uint64_t *>(c->fe_lists.p);
    w.counts = static_cast<uint32_t *>(c->fe_meta.p);
    w.join_prev = reinterpret_cast<int32_t *>(w.counts + walkers);
    w.join_next = w.join_prev + walkers;
    w.seq = static_cast<uint32_t *>(c->fe_seq.p);
    return TP_OK;
}

int enqueue_ranf_faster_evgen_jump(Ctx *c, const tp_params *params, tpk::FusedArgs &a, const tp_ranf_state *rs,
                                   uint64_t num_events, uint32_t batch_size, void *per_batch_dev)
{
    const bool f32 = params->precision == TP_PREC_F32;
    if (batch_size > 10u * TP_BATCH_SIZE)
        return fail(c, TP_ERR_ARG, "faster-evgen with jump seeding runs batches of at most 100 000 events");
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    const uint64_t chunk = std::min<uint64_t>(nb, FE_WALK_BATCHES);
    const uint64_t PS = FE_JUMP_PIECE * ((batch_size + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE);
    const uint64_t max_rounds = PS / 55 + 2;
    const int K = std::max(1, bits_for(max_rounds / RANF_R + 1));
    int rc = ensure_polys(c, RANF_R, K);
    if (!rc)
        rc = ensure(c, c->draws, chunk * PS * sizeof(int32_t));
    if (!rc)
        rc = ensure(c, c->origin, chunk * 56 * sizeof(int32_t));
    if (!rc)
        rc = ensure(c, c->p0, chunk * sizeof(int32_t));
    if (!rc)
        rc = ensure(c, c->evq, chunk * batch_size * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_flag, sizeof(uint32_t));
    if (!rc && f32)
        rc = ensure(c, c->f32ev, chunk * batch_size * 6 * sizeof(float));
    if (!rc) // each piece's bits: its segments' words, the last one cut at the piece's end
        rc = ensure_acc_words(c, (chunk + 1) * (PS / 32) + RANF_SEG_WORDS);
    if (rc)
        return rc;
    tpk::F32Args fa;
    std::memset(&fa, 0, sizeof fa);
    if (f32) {
        fa.P = make_fparams(params);
        fa.num_events = num_events;
        fa.batch_size = batch_size;
        fa.source = 2;
        fa.ev = static_cast<float *>(c->f32ev.p);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream)); // previous users of the staging copies are done
    std::memcpy(c->origin0, rs->numbers, sizeof c->origin0);
    c->p00 = static_cast<int32_t>(55u - rs->index);
    c->fe_flag_host = 0;
    HIPCHK(c, hipMemcpyAsync(c->fe_flag.p, &c->fe_flag_host, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    auto expand = [&](uint64_t first_piece, uint64_t npieces, uint64_t d_begin, uint64_t d_end) -> int {
        tpk::RanfArgs r;
        std::memset(&r, 0, sizeof r);
        r.origin = static_cast<const int32_t *>(c->origin.p) + 56 * first_piece;
        r.p0 = static_cast<const int32_t *>(c->p0.p) + first_piece;
        r.polys = static_cast<const uint32_t *>(c->polys.p);
        r.K = K;
        r.R = RANF_R;
        r.out = static_cast<int32_t *>(c->draws.p) + first_piece * PS;
        r.d_begin = d_begin;
        r.d_end = d_end;
        r.seg_first = 0;
        r.nsegs = (max_rounds + RANF_R - 1) / RANF_R + 1;
        r.piece_stride = PS;
        r.bits = f32 ? 2 : 1;
        r.bits_piece_words = PS / 32;
        r.acc1 = static_cast<uint32_t *>(c->xo_acc.p) + first_piece * (PS / 32);
        r.acc3 = r.acc1 + c->xo_acc_words;
        tpk::launch_ranf_expand(r, npieces, c->stream);
        HIPCHK(c, hipGetLastError());
        return TP_OK;
    };
    bool first = true;
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        ++c->fe_chunks;
        const uint64_t cnb = std::min(nb, b0 + chunk) - b0;
        if (first)
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        tpk::launch_ranf_seed(rs->seed, b0, cnb, static_cast<int32_t *>(c->origin.p), static_cast<int32_t *>(c->p0.p),
                              c->stream);
        HIPCHK(c, hipGetLastError());
        if (b0 == 0) {
            // batch 0 runs on the caller's state, which may be inside its
            // round: its piece is expanded from the round's start (d_begin
            // = -p0, modulo 2^64), so that buffer index = frame position
            HIPCHK(c, hipMemcpyAsync(c->origin.p, c->origin0, sizeof c->origin0, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->p0.p, &c->p00, sizeof c->p00, hipMemcpyHostToDevice, c->stream));
            const uint64_t p00 = (uint64_t)c->p00;
            rc = expand(0, 1, 0ull - p00, PS - p00);
            if (!rc && cnb > 1)
                rc = expand(1, cnb - 1, 0, PS);
        } else {
            rc = expand(0, cnb, 0, PS);
        }
        if (rc)
            return rc;
        tpk::FeBatchWalkArgs w;
        std::memset(&w, 0, sizeof w);
        rc = fe_batch_walk_setup(c, w, f32 ? 1 : 0, chunk, batch_size);
        if (rc)
            return rc;
        w.buf = static_cast<const int32_t *>(c->draws.p);
        w.acc1 = static_cast<const uint32_t *>(c->xo_acc.p);
        w.acc3 = w.acc1 + c->xo_acc_words;
        w.piece_stride = PS;
        w.p0 = static_cast<const int32_t *>(c->p0.p);
        w.nb = cnb;
        w.b_first = b0;
        w.num_events = num_events;
        w.batch_size = batch_size;
        w.evq = static_cast<uint64_t *>(c->evq.p);
        w.overflow = static_cast<uint32_t *>(c->fe_flag.p);
        tpk::launch_fe_walk_batches(w, c->stream);
        HIPCHK(c, hipGetLastError());
        if (first) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            first = false;
        }
        if (f32) {
            fa.nb = cnb;
            fa.b_first = b0;
            fa.draws = static_cast<const int32_t *>(c->draws.p);
            fa.qbase = 0;
            fa.acc1 = w.acc1;
            fa.bbase = 0;
            fa.bstride = PS;
            fa.evq = static_cast<const uint64_t *>(c->evq.p);
            fa.out = static_cast<tp_acc *>(per_batch_dev) + b0;
            tpk::launch_f32_events(fa, f32_grid(c, cnb), c->stream);
            HIPCHK(c, hipGetLastError());
            tpk::launch_f32_accumulate(fa, f32_grid(c, cnb), c->stream);
            HIPCHK(c, hipGetLastError());
            continue;
        }
        a.nb = cnb;
        a.b_first = b0;
        a.draws = static_cast<const int32_t *>(c->draws.p);
        a.qbase = 0;
        a.acc1 = w.acc1;
        a.bbase = 0;
        a.bstride = PS;
        a.evq = static_cast<const uint64_t *>(c->evq.p);
        a.out = static_cast<tp_acc *>(per_batch_dev) + b0;
        tpk::launch_fused(a, 3, grid_for(c, 3, 0, cnb), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->fe_flag_host, c->fe_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->fe_flag_host)
        return fail(c, TP_ERR_INTERNAL, "faster-evgen: a batch consumed more of its stream than was expanded");
    return TP_OK;
}

// xoshiro256+ with `faster-evgen` and jump seeding (faster-threading with
// standard-random): batch b's stream starts at J^b s0 (J the 2^128 jump,
// standard.rs:51-55) and random_array::<N> takes N consecutive draws with no
// round to discard (standard.rs:33-35).  Each chunk's batches are expanded
// into HBM as f64 uniforms with their accept bits, one piece per batch,
// walked by segment walkers, and simulated from those starts.  A batch of
// 10 000 events consumes 166 400 draws on average (16.64 per event, standard
// deviation about 204 per batch); a piece holds 172 032.

int enqueue_xo_faster_evgen_jump(Ctx *c, tpk::FusedArgs &a, const tp_xoshiro_state *x, uint64_t num_events,
                                 uint32_t batch_size, void *per_batch_dev)
{
    if (batch_size > TP_BATCH_SIZE)
        return fail(c, TP_ERR_ARG, "faster-evgen with jump seeding runs batches of at most 10 000 events");
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    const uint64_t chunk = std::min<uint64_t>(nb, FE_WALK_BATCHES);
    Tables *t = nullptr;
    int rc = ensure_tables(c, TP_SEED_JUMP, batch_size, nb, grid_for(c, 4, 0, chunk), &t);
    if (!rc)
        rc = ensure_lane_tables(c);
    if (!rc)
        rc = ensure(c, c->xo_states, chunk * 4 * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->xo_draws, chunk * FE_XO_PIECE * sizeof(double));
    if (!rc)
        rc = ensure_acc(c, chunk);
    if (!rc)
        rc = ensure(c, c->evq, chunk * batch_size * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_flag, sizeof(uint32_t));
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->fe_flag_host = 0;
    HIPCHK(c, hipMemcpyAsync(c->fe_flag.p, &c->fe_flag_host, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    tpk::XoPieceArgs p;
    std::memset(&p, 0, sizeof p);
    std::memcpy(p.s0, x->s, sizeof p.s0);
    p.bt = static_cast<const uint64_t *>(t->bt.p);
    p.kb = t->kb;
    p.lt = static_cast<const uint64_t *>(c->xo_lt.p);
    p.states = static_cast<uint64_t *>(c->xo_states.p);
    p.piece_len = FE_XO_PIECE;
    p.per_thread = FE_XO_PER_THREAD;
    p.xbuf = static_cast<double *>(c->xo_draws.p);
    p.acc1 = static_cast<uint32_t *>(c->xo_acc.p);
    p.acc3 = p.acc1 + c->xo_acc_words;
    tpk::FeBatchWalkArgs w;
    std::memset(&w, 0, sizeof w);
    rc = fe_batch_walk_setup(c, w, 2, chunk, batch_size);
    if (rc)
        return rc;
    w.acc1 = p.acc1;
    w.acc3 = p.acc3;
    w.piece_stride = FE_XO_PIECE;
    w.num_events = num_events;
    w.batch_size = batch_size;
    w.evq = static_cast<uint64_t *>(c->evq.p);
    w.overflow = static_cast<uint32_t *>(c->fe_flag.p);
    bool first = true;
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        ++c->fe_chunks;
        const uint64_t cnb = std::min(nb, b0 + chunk) - b0;
        if (first)
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        p.nb = cnb;
        p.b_first = b0;
        tpk::launch_xo_expand(p, c->stream);
        HIPCHK(c, hipGetLastError());
        w.nb = cnb;
        w.b_first = b0;
        tpk::launch_fe_walk_batches(w, c->stream);
        HIPCHK(c, hipGetLastError());
        a.nb = cnb;
        a.b_first = b0;
        a.xdraws = p.xbuf;
        a.qbase = 0;
        a.acc1 = w.acc1;
        a.bbase = 0;
        a.bstride = FE_XO_PIECE;
        a.evq = w.evq;
        a.out = static_cast<tp_acc *>(per_batch_dev) + b0;
        if (first) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            first = false;
        }
        tpk::launch_fused(a, 4, grid_for(c, 4, 0, cnb), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->fe_flag_host, c->fe_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->fe_flag_host)
        return fail(c, TP_ERR_INTERNAL, "faster-evgen: a batch consumed more of its stream than was expanded");
    return TP_OK;
}

// xoshiro128+ with `faster-evgen` (standard-random,f32,faster-evgen), the
// all-f32 mode: skip-ahead seeding walks one stream as the xoshiro256+ path
// does (f32 uniforms, the unit-disc test in f32); jump seeding (the 2^64
// jump per batch, standard.rs:51-55) gives each batch a piece of its own.
// Events go through f32_events_kernel<3> and the in-order f32 sums.
int enqueue_xo128_faster_evgen(Ctx *c, const tp_params *params, int seeding, const tp_xoshiro128_state *x,
                               uint64_t num_events, uint32_t batch_size, void *per_batch_dev)
{
    HIPCHK(c, hipStreamSynchronize(c->stream)); // the staging vectors below are reused
    uint32_t s0[4];
    std::memcpy(s0, x->s, sizeof s0);
    tpk::FusedArgs unused;
    std::memset(&unused, 0, sizeof unused);
    tpk::F32Args fa;
    std::memset(&fa, 0, sizeof fa);
    fa.P = make_fparams(params);
    fa.num_events = num_events;
    fa.batch_size = batch_size;
    fa.source = 3;
    if (seeding != TP_SEED_JUMP) {
        FeStream fs;
        fs.kind = 3;
        fs.s0_128 = s0;
        const FeHostWalk walk = [&](uint64_t q, uint64_t n, uint64_t *starts, uint64_t *q_next) {
            tp::xoshiro128_walk_faster_evgen(s0, q, n, starts, q_next);
        };
        uint64_t q_end = 0;
        int rc = enqueue_fe_chunks(c, fs, walk, unused, &fa, 0, num_events, batch_size, per_batch_dev, &q_end);
        if (rc)
            return rc;
        c->fe_xo_end = q_end;
        return TP_OK;
    }
    if (batch_size > TP_BATCH_SIZE)
        return fail(c, TP_ERR_ARG, "faster-evgen with jump seeding runs batches of at most 10 000 events");
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    const uint64_t chunk = std::min<uint64_t>(nb, FE_WALK_BATCHES);
    int rc = ensure_xo128_piece_table(c, true, FE_XO_PIECE, nb);
    if (!rc)
        rc = ensure(c, c->evq, chunk * batch_size * sizeof(uint64_t));
    if (!rc)
        rc = ensure(c, c->fe_flag, sizeof(uint32_t));
    if (!rc)
        rc = ensure(c, c->f32ev, chunk * batch_size * 6 * sizeof(float));
    if (rc)
        return rc;
    fa.ev = static_cast<float *>(c->f32ev.p);
    c->fe_flag_host = 0;
    HIPCHK(c, hipMemcpyAsync(c->fe_flag.p, &c->fe_flag_host, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    tpk::FeBatchWalkArgs w;
    std::memset(&w, 0, sizeof w);
    rc = fe_batch_walk_setup(c, w, 3, chunk, batch_size);
    if (rc)
        return rc;
    w.piece_stride = FE_XO_PIECE;
    w.num_events = num_events;
    w.batch_size = batch_size;
    w.evq = static_cast<uint64_t *>(c->evq.p);
    w.overflow = static_cast<uint32_t *>(c->fe_flag.p);
    bool first = true;
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        ++c->fe_chunks;
        const uint64_t cnb = std::min(nb, b0 + chunk) - b0;
        if (first)
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        rc = xo128_expand(c, s0, b0, cnb);
        if (rc)
            return rc;
        w.acc1 = static_cast<const uint32_t *>(c->xo_acc.p);
        w.acc3 = w.acc1 + c->xo_acc_words;
        w.nb = cnb;
        w.b_first = b0;
        tpk::launch_fe_walk_batches(w, c->stream);
        HIPCHK(c, hipGetLastError());
        if (first) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            first = false;
        }
        fa.nb = cnb;
        fa.b_first = b0;
        fa.x32draws = static_cast<const float *>(c->xo128_draws.p);
        fa.qbase = 0;
        fa.acc1 = w.acc1;
        fa.bbase = 0;
        fa.bstride = FE_XO_PIECE;
        fa.evq = w.evq;
        fa.out = static_cast<tp_acc *>(per_batch_dev) + b0;
        tpk::launch_f32_events(fa, f32_grid(c, cnb), c->stream);
        HIPCHK(c, hipGetLastError());
        tpk::launch_f32_accumulate(fa, f32_grid(c, cnb), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->fe_flag_host, c->fe_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->fe_flag_host)
        return fail(c, TP_ERR_INTERNAL, "faster-evgen: a batch consumed more of its stream than was expanded");
    return TP_OK;
}

// Enqueue a whole run (or one closure-sized batch when batch_size == n).
int enqueue_run(Ctx *c, const tp_params *params, int rng_kind, int seeding, const void *state0,
                uint64_t num_events, uint32_t batch_size, void *per_batch_dev)
{
    const uint64_t nb = (num_events + batch_size - 1) / batch_size;
    tpk::FusedArgs a;
    std::memset(&a, 0, sizeof a);
    a.P = make_params(params);
    a.num_events = num_events;
    a.batch_size = batch_size;
    a.precision = params->precision;
    int rc = ensure_mt(c);
    if (rc)
        return rc;
    a.mt = static_cast<const uint64_t *>(c->mt.p);
    a.km = c->km;
    if (rng_kind == TP_RNG_XOSHIRO128 && params->faster_evgen)
        return enqueue_xo128_faster_evgen(c, params, seeding, static_cast<const tp_xoshiro128_state *>(state0),
                                          num_events, batch_size, per_batch_dev);
    if (rng_kind == TP_RNG_XOSHIRO128)
        return enqueue_xo128_f32(c, params, seeding, static_cast<const tp_xoshiro128_state *>(state0), num_events,
                                 batch_size, per_batch_dev);
    const bool f32 = params->precision == TP_PREC_F32;
    tpk::F32Args fa;
    std::memset(&fa, 0, sizeof fa);
    if (f32) {
        fa.P = make_fparams(params);
        fa.num_events = num_events;
        fa.batch_size = batch_size;
        fa.source = 1;
        rc = ensure(c, c->f32ev, std::min<uint64_t>(nb, RANF_CHUNK_BATCHES) * batch_size * 6 * sizeof(float));
        if (rc)
            return rc;
        fa.ev = static_cast<float *>(c->f32ev.p);
    }
    if (rng_kind == TP_RNG_XOSHIRO && params->faster_evgen) {
        if (seeding == TP_SEED_JUMP)
            return enqueue_xo_faster_evgen_jump(c, a, static_cast<const tp_xoshiro_state *>(state0), num_events,
                                                batch_size, per_batch_dev);
        return enqueue_xo_faster_evgen(c, a, static_cast<const tp_xoshiro_state *>(state0), num_events, batch_size,
                                       per_batch_dev);
    }
    if (rng_kind == TP_RNG_XOSHIRO) {
        const tp_xoshiro_state *x = static_cast<const tp_xoshiro_state *>(state0);
        std::memcpy(a.s0, x->s, sizeof a.s0);
        const int grid = grid_for(c, 0, a.precision, nb);
        Tables *t = nullptr;
        rc = ensure_tables(c, seeding, batch_size, nb, grid, &t);
        if (rc)
            return rc;
        a.bt = static_cast<const uint64_t *>(t->bt.p);
        a.stride = static_cast<const uint64_t *>(t->stride.p);
        a.kb = t->kb;
        a.nb = nb;
        a.b_first = 0;
        a.out = per_batch_dev;
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        tpk::launch_fused(a, 0, grid, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        return TP_OK;
    }
    // RANF: expand the stream of a chunk of batches into HBM, then run the
    // fused kernel on it.
    const tp_ranf_state *rs = static_cast<const tp_ranf_state *>(state0);
    if (rs->index > 55)
        return fail(c, TP_ERR_ARG, "RANF state index out of range");
    if (params->faster_evgen) {
        if (seeding == TP_SEED_JUMP)
            return enqueue_ranf_faster_evgen_jump(c, params, a, rs, num_events, batch_size, per_batch_dev);
        if (f32)
            fa.source = 2;
        return enqueue_ranf_faster_evgen(c, a, f32 ? &fa : nullptr, rs, num_events, batch_size, per_batch_dev);
    }
    const uint64_t chunk = std::min<uint64_t>(nb, RANF_CHUNK_BATCHES);
    const size_t per_batch_draws = 12ull * batch_size;
    rc = ensure(c, c->draws, chunk * per_batch_draws * sizeof(int32_t));
    if (rc)
        return rc;
    rc = ensure(c, c->origin, chunk * 56 * sizeof(int32_t));
    if (rc)
        return rc;
    rc = ensure(c, c->p0, chunk * sizeof(int32_t));
    if (rc)
        return rc;
    // Segments needed to cover one chunk: every segment spans R rounds.
    const uint64_t total_draws = 12ull * num_events;
    const uint64_t p0_master = 55u - rs->index;
    uint64_t max_rounds = (seeding == TP_SEED_JUMP) ? (per_batch_draws + 54) / 55 + 1
                                                    : (p0_master + total_draws + 54) / 55 + 1;
    int K = std::max(1, bits_for(max_rounds / RANF_R + 1));
    rc = ensure_polys(c, RANF_R, K);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream)); // previous users of the staging copy are done
    int32_t *origin0 = c->origin0;
    std::memcpy(origin0, rs->numbers, sizeof c->origin0);
    c->p00 = static_cast<int32_t>(p0_master);
    int32_t &p00 = c->p00;
    bool first_kernel = true;
    for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
        const uint64_t b1 = std::min(nb, b0 + chunk);
        const uint64_t cnb = b1 - b0;
        tpk::RanfArgs r;
        std::memset(&r, 0, sizeof r);
        r.origin = static_cast<const int32_t *>(c->origin.p);
        r.p0 = static_cast<const int32_t *>(c->p0.p);
        r.polys = static_cast<const uint32_t *>(c->polys.p);
        r.K = K;
        r.R = RANF_R;
        r.out = static_cast<int32_t *>(c->draws.p);
        uint64_t npieces;
        if (seeding == TP_SEED_JUMP) {
            tpk::launch_ranf_seed(rs->seed, b0, cnb, static_cast<int32_t *>(c->origin.p),
                                  static_cast<int32_t *>(c->p0.p), c->stream);
            HIPCHK(c, hipGetLastError());
            if (b0 == 0) { // batch 0 runs on the caller's state itself
                HIPCHK(c, hipMemcpyAsync(c->origin.p, origin0, sizeof c->origin0, hipMemcpyHostToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(c->p0.p, &p00, sizeof p00, hipMemcpyHostToDevice, c->stream));
            }
            npieces = cnb;
            r.d_begin = 0;
            r.d_end = per_batch_draws;
            r.seg_first = 0;
            r.piece_stride = per_batch_draws;
            r.nsegs = ((uint64_t)(max_rounds) + RANF_R - 1) / RANF_R + 1;
        } else {
            if (b0 == 0) {
                HIPCHK(c, hipMemcpyAsync(c->origin.p, origin0, sizeof c->origin0, hipMemcpyHostToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(c->p0.p, &p00, sizeof p00, hipMemcpyHostToDevice, c->stream));
            }
            npieces = 1;
            r.d_begin = b0 * per_batch_draws;
            r.d_end = std::min<uint64_t>(b1 * per_batch_draws, total_draws);
            const uint64_t q_begin = p0_master + r.d_begin, q_end = p0_master + r.d_end;
            r.seg_first = q_begin / 55 / RANF_R;
            const uint64_t seg_last = (q_end - 1) / 55 / RANF_R;
            r.nsegs = seg_last - r.seg_first + 1;
            r.piece_stride = 0;
        }
        if (first_kernel) {
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        }
        tpk::launch_ranf_expand(r, npieces, c->stream);
        HIPCHK(c, hipGetLastError());
        a.nb = cnb;
        a.b_first = b0;
        a.draws = static_cast<const int32_t *>(c->draws.p);
        a.out = static_cast<tp_acc *>(per_batch_dev) + b0;
        if (first_kernel) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            first_kernel = false;
        }
        if (f32) {
            fa.nb = cnb;
            fa.b_first = b0;
            fa.draws = a.draws;
            fa.out = a.out;
            tpk::launch_f32_events(fa, f32_grid(c, cnb), c->stream);
            HIPCHK(c, hipGetLastError());
            tpk::launch_f32_accumulate(fa, f32_grid(c, cnb), c->stream);
        } else {
            tpk::launch_fused(a, 1, grid_for(c, 1, a.precision, cnb), c->stream);
        }
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    return TP_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
        return 0.f;
    return ms;
}

} // namespace

// ---- tables of the no-allocation entry point ---------------------------------
// tp_run_device_noalloc runs the xoshiro256+ event loop without allocating
// device memory: the GF(2) tables its kernels read live in `scratch`, a
// file-scope device array, filled once per device at set-up (tp_create) and
// only read after that, so calls on one device that overlap do not race on
// it.  Each device holds its own copy at its own address, which set-up takes
// with hipGetSymbolAddress while that device is current; host code uses that
// address, never the array's name.
//   matrices [0, KM)                M^(2^k), the in-batch draw offsets
//   [KM, KM + NS_KB)                B^(2^k), B = M^(12 * 10 000) (skip-ahead)
//   [KM + NS_KB, KM + 2 NS_KB)      B^(2^k), B = the 2^128 jump (jump seeding)
//   KM + 2 NS_KB + 2 seeding + p    B^grid of the full persistent grid, p = 0 f64, 1 mixed
constexpr int NS_KB = 32;                            // runs of fewer than 2^32 batches
constexpr int NS_BT = KM, NS_STRIDE = KM + 2 * NS_KB, NS_MATS = KM + 2 * NS_KB + 4;
constexpr size_t NS_WORDS = (size_t)NS_MATS * 1024; // 100 matrices of 8 KiB: 800 KiB
static_assert(sizeof(tp::Gf2Mat) == 1024 * sizeof(uint64_t), "matrix layout");
static __device__ __align__(256) uint64_t scratch[NS_WORDS];

namespace {
constexpr int NS_MAX_DEVICES = 64;
struct NoallocSetup {
    std::once_flag once;
    int rc = TP_ERR_INTERNAL;
    std::string err;
    uint64_t *ptr = nullptr; // this device's address of `scratch`
    int full_grid[2] = {0, 0};
};
NoallocSetup g_noalloc[NS_MAX_DEVICES];

// Once per device, with the device current: the array's address and its tables.
void noalloc_setup(NoallocSetup &S, int cus, const int blocks[2])
{
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        S.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        S.rc = TP_ERR_HIP;
        return;
    }
    std::vector<tp::Gf2Mat> mats(NS_MATS);
    const std::vector<tp::Gf2Mat> &mt = tp::xoshiro_step_pow2();
    for (int k = 0; k < KM; ++k)
        mats[k] = mt[k];
    tp::Gf2Mat B[2];
    tp::gf2_pow_from_table(mt, 12ull * TP_BATCH_SIZE, B[0]);
    tp::gf2_jump128(B[1]);
    for (int sd = 0; sd < 2; ++sd) {
        const std::vector<tp::Gf2Mat> bt = tp::gf2_pow2_table(B[sd], NS_KB);
        for (int k = 0; k < NS_KB; ++k)
            mats[NS_BT + sd * NS_KB + k] = bt[k];
        for (int p = 0; p < 2; ++p) {
            S.full_grid[p] = std::max(1, cus * blocks[p]);
            tp::gf2_pow_from_table(bt, (uint64_t)S.full_grid[p], mats[NS_STRIDE + 2 * sd + p]);
        }
    }
    e = hipMemcpy(ptr, mats.data(), NS_WORDS * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        S.err = std::string("scratch upload: ") + hipGetErrorString(e);
        S.rc = TP_ERR_HIP;
        return;
    }
    S.ptr = static_cast<uint64_t *>(ptr);
    S.rc = TP_OK;
}
} // namespace

extern "C" int tp_create(int device, void **ctx_out)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!ctx_out)
            return TP_ERR_ARG;
        *ctx_out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            return TP_ERR_HIP;
        if (device < 0 || device >= count)
            return TP_ERR_ARG;
        if (hipSetDevice(device) != hipSuccess)
            return TP_ERR_HIP;
        Ctx *c = new Ctx;
        c->device = device;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return TP_ERR_HIP;
        }
        for (auto &e : c->ev)
            if (hipEventCreate(&e) != hipSuccess) {
                delete c;
                return TP_ERR_HIP;
            }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->cus = prop.multiProcessorCount;
        for (int src = 0; src < 5; ++src)
            for (int prec = 0; prec < 2; ++prec)
                c->blocks_per_cu[src][prec] = tpk::fused_blocks_per_cu(src, prec);
        if (device < NS_MAX_DEVICES) { // tables of tp_run_device_noalloc (device current)
            NoallocSetup &S = g_noalloc[device];
            std::call_once(S.once, [&] { noalloc_setup(S, c->cus, c->blocks_per_cu[0]); });
        }
        *ctx_out = c;
        return TP_OK;
    });
}

extern "C" int tp_destroy(void *ctx)
{
    return tp_guard(nullptr, [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (DevBuf *b : {&c->mt, &c->polys, &c->draws, &c->origin, &c->p0, &c->per_batch, &c->merged, &c->evq,
                          &c->mt128, &c->bt128, &c->st128, &c->f32ev, &c->fe_lists, &c->fe_meta, &c->fe_piece,
                          &c->fe_flag, &c->fe_seq, &c->xo_states, &c->xo_draws, &c->xo128_states,
                          &c->xo128_draws, &c->xo128_pt, &c->xo_acc, &c->xo_lt, &c->xo128_lt, &c->fe_stat, &c->adj_u})
            if (b->p)
                (void)hipFree(b->p);
        for (auto &kv : c->tables) {
            if (kv.second->bt.p)
                (void)hipFree(kv.second->bt.p);
            if (kv.second->stride.p)
                (void)hipFree(kv.second->stride.p);
        }
        for (auto &e : c->ev)
            if (e)
                (void)hipEventDestroy(e);
        if (c->fe_stat_host)
            (void)hipHostFree(c->fe_stat_host);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return TP_OK;
    });
}

extern "C" const char *tp_last_error(const void *ctx)
{
    const Ctx *c = static_cast<const Ctx *>(ctx);
    return c ? c->err.c_str() : "null context";
}

extern "C" int tp_rng_seed(int rng_kind, uint64_t seed, void *state)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!state)
            return TP_ERR_ARG;
        if (rng_kind == TP_RNG_RANF) {
            tp::RanfState g;
            tp::ranf_seeded(g, static_cast<int32_t>(seed));
            tp_ranf_state *o = static_cast<tp_ranf_state *>(state);
            std::memcpy(o->numbers, g.numbers, sizeof o->numbers);
            o->index = g.index;
            o->seed = g.seed;
            return TP_OK;
        }
        if (rng_kind == TP_RNG_XOSHIRO) {
            tp::xoshiro_seed(seed, static_cast<tp_xoshiro_state *>(state)->s);
            return TP_OK;
        }
        if (rng_kind == TP_RNG_XOSHIRO128) {
            tp::xoshiro128_seed(seed, static_cast<tp_xoshiro128_state *>(state)->s);
            return TP_OK;
        }
        return TP_ERR_ARG;
    });
}

extern "C" int tp_rng_init(int rng_kind, void *state)
{
    return tp_guard(nullptr, [&]() -> int {
        return tp_rng_seed(rng_kind, rng_kind == TP_RNG_RANF ? static_cast<uint64_t>(tp::RANF_SEED) : 12345u, state);
    });
}

extern "C" int tp_rng_skip(int rng_kind, void *state, uint64_t n)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!state)
            return TP_ERR_ARG;
        if (rng_kind == TP_RNG_XOSHIRO) {
            tp::xoshiro_advance(static_cast<tp_xoshiro_state *>(state)->s, n);
            return TP_OK;
        }
        if (rng_kind == TP_RNG_XOSHIRO128) {
            tp::xoshiro128_advance(static_cast<tp_xoshiro128_state *>(state)->s, n);
            return TP_OK;
        }
        if (rng_kind != TP_RNG_RANF)
            return TP_ERR_ARG;
        tp_ranf_state *s = static_cast<tp_ranf_state *>(state);
        if (s->index > 55)
            return TP_ERR_ARG;
        tp::RanfState g;
        std::memcpy(g.numbers, s->numbers, sizeof g.numbers);
        g.index = s->index;
        g.seed = s->seed;
        tp::ranf_advance(g, n);
        std::memcpy(s->numbers, g.numbers, sizeof g.numbers);
        s->index = g.index;
        return TP_OK;
    });
}

extern "C" int tp_rng_jump(int rng_kind, void *state)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!state)
            return TP_ERR_ARG;
        if (rng_kind == TP_RNG_XOSHIRO) {
            static tp::Gf2Mat J;
            static std::once_flag once;
            std::call_once(once, [] { tp::gf2_jump128(J); });
            uint64_t *s = static_cast<tp_xoshiro_state *>(state)->s;
            tp::gf2_apply(J, s, s);
            return TP_OK;
        }
        if (rng_kind == TP_RNG_XOSHIRO128) {
            tp::xoshiro128_jump(static_cast<tp_xoshiro128_state *>(state)->s);
            return TP_OK;
        }
        if (rng_kind != TP_RNG_RANF)
            return TP_ERR_ARG;
        tp_ranf_state *s = static_cast<tp_ranf_state *>(state);
        const int32_t seed = static_cast<int32_t>(static_cast<uint32_t>(s->seed) + 123456u);
        return tp_rng_seed(TP_RNG_RANF, static_cast<uint64_t>(static_cast<uint32_t>(seed)), state);
    });
}

extern "C" int tp_events(void *ctx, const tp_params *params, int rng_kind, const void *rng_state, uint32_t n,
                         double *u, double *p, int32_t *keep, double *m)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        int rc = check_common(c, params, rng_kind, rng_state);
        if (rc)
            return rc;
        if (n == 0)
            return TP_OK;
        if (n > (1u << 20))
            return fail(c, TP_ERR_ARG, "tp_events: n exceeds 2^20");
        if (params->faster_evgen || params->precision == TP_PREC_F32)
            return fail(c, TP_ERR_UNSUPPORTED, "tp_events: per-event values of faster-evgen and all-f32 runs are not available");
        HIPCHK(c, hipSetDevice(c->device));
        tpk::FusedArgs a;
        std::memset(&a, 0, sizeof a);
        a.P = make_params(params);
        a.precision = params->precision;
        rc = ensure_mt(c);
        if (rc)
            return rc;
        a.mt = static_cast<const uint64_t *>(c->mt.p);
        a.km = c->km;
        int source = 0;
        if (rng_kind == TP_RNG_XOSHIRO) {
            std::memcpy(a.s0, static_cast<const tp_xoshiro_state *>(rng_state)->s, sizeof a.s0);
        } else {
            // Expand 12 n draws from the given state: one piece, skip-ahead.
            const tp_ranf_state *rs = static_cast<const tp_ranf_state *>(rng_state);
            if (rs->index > 55)
                return fail(c, TP_ERR_ARG, "RANF state index out of range");
            const uint64_t D = 12ull * n;
            rc = ensure(c, c->draws, D * sizeof(int32_t));
            if (!rc)
                rc = ensure(c, c->origin, 56 * sizeof(int32_t));
            if (!rc)
                rc = ensure(c, c->p0, sizeof(int32_t));
            if (rc)
                return rc;
            HIPCHK(c, hipStreamSynchronize(c->stream));
            std::memcpy(c->origin0, rs->numbers, sizeof c->origin0);
            c->p00 = static_cast<int32_t>(55u - rs->index);
            const uint64_t max_rounds = (c->p00 + D + 54) / 55 + 1;
            const int K = std::max(1, bits_for(max_rounds / RANF_R + 1));
            rc = ensure_polys(c, RANF_R, K);
            if (rc)
                return rc;
            HIPCHK(c, hipMemcpyAsync(c->origin.p, c->origin0, sizeof c->origin0, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->p0.p, &c->p00, sizeof c->p00, hipMemcpyHostToDevice, c->stream));
            tpk::RanfArgs r;
            std::memset(&r, 0, sizeof r);
            r.origin = static_cast<const int32_t *>(c->origin.p);
            r.p0 = static_cast<const int32_t *>(c->p0.p);
            r.polys = static_cast<const uint32_t *>(c->polys.p);
            r.K = K;
            r.R = RANF_R;
            r.out = static_cast<int32_t *>(c->draws.p);
            r.d_begin = 0;
            r.d_end = D;
            r.seg_first = 0;
            r.nsegs = (c->p00 + D - 1) / 55 / RANF_R + 1;
            tpk::launch_ranf_expand(r, 1, c->stream);
            HIPCHK(c, hipGetLastError());
            a.draws = static_cast<const int32_t *>(c->draws.p);
            source = 1;
        }
        ScopedBufs<4> bufs; // freed on every return below
        DevBuf &bu = bufs[0], &bp = bufs[1], &bk = bufs[2], &bm = bufs[3];
        tpk::EventOut o{nullptr, nullptr, nullptr, nullptr};
        struct Item {
            void *host;
            DevBuf *buf;
            size_t bytes;
        } items[4] = {{u, &bu, 12ull * n * sizeof(double)},
                      {p, &bp, 12ull * n * sizeof(double)},
                      {keep, &bk, n * sizeof(int32_t)},
                      {m, &bm, 5ull * n * sizeof(double)}};
        for (Item &it : items)
            if (it.host && (rc = ensure(c, *it.buf, it.bytes)) != TP_OK)
                break;
        if (rc == TP_OK) {
            o.u = static_cast<double *>(bu.p);
            o.p = static_cast<double *>(bp.p);
            o.keep = static_cast<int32_t *>(bk.p);
            o.m = static_cast<double *>(bm.p);
            tpk::launch_events(a, source, n, o, c->stream);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                rc = fail(c, TP_ERR_HIP, std::string("events kernel: ") + hipGetErrorString(e));
            for (Item &it : items)
                if (rc == TP_OK && it.host) {
                    e = hipMemcpyAsync(it.host, it.buf->p, it.bytes, hipMemcpyDeviceToHost, c->stream);
                    if (e != hipSuccess)
                        rc = fail(c, TP_ERR_HIP, std::string("events copy: ") + hipGetErrorString(e));
                }
            e = hipStreamSynchronize(c->stream);
            if (rc == TP_OK && e != hipSuccess)
                rc = fail(c, TP_ERR_HIP, std::string("events sync: ") + hipGetErrorString(e));
        }
        return rc;
    });
}

extern "C" int tp_adjoint(void *ctx, const tp_params *params, int rng_kind, const void *rng_state, const void *u_in,
                          uint64_t n, int precision, int on_device, void *F, void *dF_du, void *dF_dp)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        if (!params)
            return fail(c, TP_ERR_ARG, "null params");
        if (precision != TP_PREC_F64 && precision != TP_PREC_F32)
            return fail(c, TP_ERR_ARG, "tp_adjoint: precision must be TP_PREC_F64 or TP_PREC_F32");
        if (!u_in && rng_kind != TP_RNG_XOSHIRO)
            return fail(c, TP_ERR_UNSUPPORTED, "tp_adjoint draws its points from a xoshiro stream or u_in");
        if (!u_in && !rng_state)
            return fail(c, TP_ERR_ARG, "null rng state");
        if (n == 0)
            return TP_OK;
        if (!u_in && 12 * n >= (1ull << KM))
            return fail(c, TP_ERR_ARG, "tp_adjoint: n exceeds the stream offset range");
        HIPCHK(c, hipSetDevice(c->device));
        int rc = ensure_mt(c);
        if (rc)
            return rc;
        tpk::AdjArgs a;
        std::memset(&a, 0, sizeof a);
        a.P = make_params(params);
        a.w = params->ev_weight;
        a.n = n;
        a.mt = static_cast<const uint64_t *>(c->mt.p);
        a.km = c->km;
        if (!u_in) {
            std::memcpy(a.s0, static_cast<const tp_xoshiro_state *>(rng_state)->s, sizeof a.s0);
            rc = ensure(c, c->adj_u, 12 * n * sizeof(double)); // the stream's points, expanded first
            if (rc)
                return rc;
            a.u_gen = static_cast<double *>(c->adj_u.p);
        }
        const size_t es = precision == TP_PREC_F32 ? sizeof(float) : sizeof(double);
        const size_t bytes[3] = {5 * n * es, 60 * n * es, 60 * n * es};
        void *outs[3] = {F, dF_du, dF_dp};
        ScopedBufs<4> tmp; // freed on every return below
        if (on_device) {
            a.u_in = static_cast<const double *>(u_in);
            a.F = F;
            a.dF_du = dF_du;
            a.dF_dp = dF_dp;
        } else {
            if (u_in) {
                rc = ensure(c, tmp[3], 12 * n * sizeof(double));
                if (rc)
                    return rc;
                HIPCHK(c, hipMemcpyAsync(tmp[3].p, u_in, 12 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
                a.u_in = static_cast<const double *>(tmp[3].p);
            }
            for (int k = 0; k < 3 && rc == TP_OK; ++k)
                if (outs[k])
                    rc = ensure(c, tmp[k], bytes[k]);
            a.F = F ? tmp[0].p : nullptr;
            a.dF_du = dF_du ? tmp[1].p : nullptr;
            a.dF_dp = dF_dp ? tmp[2].p : nullptr;
        }
        if (rc == TP_OK) {
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            tpk::launch_adjoint(a, precision, c->stream);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                rc = fail(c, TP_ERR_HIP, std::string("adjoint kernel: ") + hipGetErrorString(e));
            HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        }
        if (rc == TP_OK && !on_device)
            for (int k = 0; k < 3 && rc == TP_OK; ++k)
                if (outs[k]) {
                    hipError_t e = hipMemcpyAsync(outs[k], tmp[k].p, bytes[k], hipMemcpyDeviceToHost, c->stream);
                    if (e != hipSuccess)
                        rc = fail(c, TP_ERR_HIP, std::string("adjoint copy: ") + hipGetErrorString(e));
                }
        hipError_t e = hipStreamSynchronize(c->stream);
        if (rc == TP_OK && e != hipSuccess)
            rc = fail(c, TP_ERR_HIP, std::string("adjoint sync: ") + hipGetErrorString(e));
        if (rc == TP_OK) {
            c->last.kernel_ms = elapsed(c->ev[1], c->ev[2]);
            c->last.device_ms = c->last.kernel_ms;
            c->last.events = n;
            c->last.batches = 0;
            c->last.host_walk_chunks = 0;
        c->last.chunks = 0;
            c->last.chunks = 0;
        }
        return rc;
    });
}

extern "C" int tp_simulate_batch(void *ctx, const tp_params *params, int rng_kind, void *rng_state, uint32_t n,
                                 tp_acc *out)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        int rc = check_common(c, params, rng_kind, rng_state);
        if (rc)
            return rc;
        if (!out)
            return fail(c, TP_ERR_ARG, "null output");
        if (n > 64u * TP_BATCH_SIZE)
            return fail(c, TP_ERR_ARG, "n exceeds 640000 events per call");
        HIPCHK(c, hipSetDevice(c->device));
        if (n == 0) { // the reference's closure returns an empty accumulator
            std::memset(out, 0, sizeof *out);
            return TP_OK;
        }
        rc = ensure(c, c->per_batch, sizeof(tp_acc));
        if (rc)
            return rc;
        rc = enqueue_run(c, params, rng_kind, TP_SEED_SKIPAHEAD, rng_state, n, n, c->per_batch.p);
        if (rc)
            return rc;
        HIPCHK(c, hipMemcpyAsync(out, c->per_batch.p, sizeof(tp_acc), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // `&mut rng` side effect: 12 draws per event (evgen.rs:182-187), or the
        // walk of the faster-evgen draws.
        if (params->faster_evgen && rng_kind == TP_RNG_XOSHIRO) {
            tp::xoshiro_advance(static_cast<tp_xoshiro_state *>(rng_state)->s, c->fe_xo_end);
        } else if (params->faster_evgen && rng_kind == TP_RNG_XOSHIRO128) {
            tp::xoshiro128_advance(static_cast<tp_xoshiro128_state *>(rng_state)->s, c->fe_xo_end);
        } else if (params->faster_evgen) {
            tp_ranf_state *s = static_cast<tp_ranf_state *>(rng_state);
            std::memcpy(s->numbers, c->fe_end.numbers, sizeof s->numbers);
            s->index = c->fe_end.index;
        } else if (rng_kind == TP_RNG_XOSHIRO) {
            tp::xoshiro_advance(static_cast<tp_xoshiro_state *>(rng_state)->s, 12ull * n);
        } else if (rng_kind == TP_RNG_XOSHIRO128) {
            tp::xoshiro128_advance(static_cast<tp_xoshiro128_state *>(rng_state)->s, 12ull * n);
        } else {
            tp_ranf_state *s = static_cast<tp_ranf_state *>(rng_state);
            tp::RanfState g;
            std::memcpy(g.numbers, s->numbers, sizeof g.numbers);
            g.index = s->index;
            g.seed = s->seed;
            tp::ranf_advance(g, 12ull * n);
            std::memcpy(s->numbers, g.numbers, sizeof g.numbers);
            s->index = g.index;
        }
        return TP_OK;
    });
}

extern "C" int tp_run_device(void *ctx, const tp_params *params, int rng_kind, int seeding, const void *rng_state0,
                             uint64_t num_events, tp_acc *per_batch_dev, tp_acc *merged_dev, tp_stats *stats)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        int rc = check_common(c, params, rng_kind, rng_state0);
        if (rc)
            return rc;
        if (seeding != TP_SEED_SKIPAHEAD && seeding != TP_SEED_JUMP)
            return fail(c, TP_ERR_ARG, "unknown seeding " + std::to_string(seeding));
        if (num_events == 0)
            return fail(c, TP_ERR_ARG, "Must simulate at least one event");
        if (!per_batch_dev)
            return fail(c, TP_ERR_ARG, "null per-batch output");
        HIPCHK(c, hipSetDevice(c->device));
        const uint64_t nb = (num_events + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE;
        // Event 0 and 1 brackets the first kernel; 2 follows the event loop.
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        c->fe_host_chunks = 0;
        c->fe_chunks = 0;
        rc = enqueue_run(c, params, rng_kind, seeding, rng_state0, num_events, TP_BATCH_SIZE, per_batch_dev);
        if (rc)
            return rc;
        if (merged_dev) {
            tpk::launch_merge(per_batch_dev, nb, merged_dev, c->stream, params->precision == TP_PREC_F32);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
        c->last.events = num_events;
        c->last.batches = nb;
        c->last.host_walk_chunks = c->fe_host_chunks;
        c->last.chunks = c->fe_chunks;
        c->timed = true;
        if (stats)
            *stats = c->last;
        return TP_OK;
    });
}

extern "C" int tp_run_device_noalloc(void *ctx, const tp_params *params, int rng_kind, int seeding,
                                     const void *rng_state0, uint64_t num_events, tp_acc *per_batch_dev,
                                     tp_acc *merged_dev, tp_stats *stats)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        int rc = check_common(c, params, rng_kind, rng_state0);
        if (rc)
            return rc;
        if (rng_kind != TP_RNG_XOSHIRO || params->faster_evgen ||
            (params->precision != TP_PREC_F64 && params->precision != TP_PREC_MIXED))
            return fail(c, TP_ERR_UNSUPPORTED,
                        "tp_run_device_noalloc runs the xoshiro256+ stream in f64 or mixed precision only");
        if (seeding != TP_SEED_SKIPAHEAD && seeding != TP_SEED_JUMP)
            return fail(c, TP_ERR_ARG, "unknown seeding");
        if (num_events == 0)
            return fail(c, TP_ERR_ARG, "Must simulate at least one event");
        if (!per_batch_dev)
            return fail(c, TP_ERR_ARG, "null per-batch output");
        const uint64_t nb = (num_events + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE;
        if (nb >= (1ull << NS_KB))
            return fail(c, TP_ERR_ARG, "tp_run_device_noalloc: at most 2^32 - 1 batches of 10 000 events");
        if (c->device >= NS_MAX_DEVICES)
            return fail(c, TP_ERR_UNSUPPORTED, "tp_run_device_noalloc: device ordinal above 63");
        const NoallocSetup &S = g_noalloc[c->device];
        if (S.rc != TP_OK)
            return fail(c, S.rc, "tp_run_device_noalloc set-up failed: " + S.err);
        HIPCHK(c, hipSetDevice(c->device));
        const int p = params->precision == TP_PREC_MIXED ? 1 : 0;
        const int grid = grid_for(c, 0, params->precision, nb);
        if ((uint64_t)grid < nb && grid != S.full_grid[p])
            return fail(c, TP_ERR_INTERNAL, "tp_run_device_noalloc: grid differs from the set-up's");
        tpk::FusedArgs a;
        std::memset(&a, 0, sizeof a);
        a.P = make_params(params);
        a.num_events = num_events;
        a.batch_size = TP_BATCH_SIZE;
        a.precision = params->precision;
        std::memcpy(a.s0, static_cast<const tp_xoshiro_state *>(rng_state0)->s, sizeof a.s0);
        a.mt = S.ptr;
        a.km = KM;
        a.bt = S.ptr + (size_t)(NS_BT + seeding * NS_KB) * 1024;
        a.kb = std::max(1, bits_for(nb));
        a.stride = S.ptr + (size_t)(NS_STRIDE + 2 * seeding + p) * 1024;
        a.nb = nb;
        a.b_first = 0;
        a.out = per_batch_dev;
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        tpk::launch_fused(a, 0, grid, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        if (merged_dev) {
            tpk::launch_merge(per_batch_dev, nb, merged_dev, c->stream, 0);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
        c->last.events = num_events;
        c->last.batches = nb;
        c->last.host_walk_chunks = 0;
        c->last.chunks = 0;
        c->timed = true;
        if (stats)
            *stats = c->last;
        return TP_OK;
    });
}

extern "C" int tp_sync(void *ctx, tp_stats *stats)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->timed) {
            c->last.device_ms = elapsed(c->ev[0], c->ev[3]);
            c->last.kernel_ms = elapsed(c->ev[1], c->ev[2]);
        }
        if (stats)
            *stats = c->last;
        return TP_OK;
    });
}

extern "C" int tp_run(void *ctx, const tp_params *params, int rng_kind, int seeding, const void *rng_state0,
                      uint64_t num_events, tp_acc *per_batch_out, tp_acc *merged_out, tp_stats *stats)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c)
            return TP_ERR_ARG;
        if (!per_batch_out)
            return fail(c, TP_ERR_ARG, "null per-batch output");
        if (num_events == 0)
            return fail(c, TP_ERR_ARG, "Must simulate at least one event");
        const uint64_t nb = (num_events + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE;
        int rc = ensure(c, c->per_batch, nb * sizeof(tp_acc));
        if (rc)
            return rc;
        rc = tp_run_device(ctx, params, rng_kind, seeding, rng_state0, num_events, static_cast<tp_acc *>(c->per_batch.p),
                           nullptr, nullptr);
        if (rc)
            return rc;
        HIPCHK(c, hipMemcpyAsync(per_batch_out, c->per_batch.p, nb * sizeof(tp_acc), hipMemcpyDeviceToHost, c->stream));
        rc = tp_sync(ctx, stats);
        if (rc)
            return rc;
        if (merged_out && params->precision == TP_PREC_F32) {
            tp_acc m = per_batch_out[0];
            for (uint64_t b = 1; b < nb; ++b)
                tp_acc_merge_f32(&m, &per_batch_out[b]);
            *merged_out = m;
        } else if (merged_out) { // ReproducibleAccumulator::get_merged_result (multi_threading.rs:107-126)
            tp_acc m = per_batch_out[0];
            for (uint64_t b = 1; b < nb; ++b) {
                m.selected_events += per_batch_out[b].selected_events;
                for (int i = 0; i < 5; ++i)
                    m.spm2[i] += per_batch_out[b].spm2[i];
                for (int i = 0; i < 5; ++i)
                    m.vars[i] += per_batch_out[b].vars[i];
                m.sigma += per_batch_out[b].sigma;
                m.variance += per_batch_out[b].variance;
            }
            *merged_out = m;
        }
        return TP_OK;
    });
}
