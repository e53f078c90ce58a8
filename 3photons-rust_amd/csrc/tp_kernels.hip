// tp_kernels.hip -- gfx950 kernels of the 3photons event loop.
//
//  fused_kernel      one wavefront per 10 000-event batch, grid-strided over
//                    batches; uniforms from xoshiro256+ in registers or from
//                    the expanded RANF stream, RAMBO -> cuts in registers,
//                    passing events compacted through an LDS queue, spinors
//                    -> amplitudes -> weight on full wavefronts, 12 running
//                    sums per lane, butterfly reduction to one accumulator
//                    per batch.
//  ranf_expand       RANF stream into HBM, one wavefront per segment of R
//                    rounds, 55 lanes = the 55 values of a round.
//  ranf_seed         jump-mode origins, seeded_new(seed + 123456 b) per batch.
//  merge_kernel      in-order left fold of the batch accumulators.
//
// Compiled with -ffp-contract=off: the f64 path keeps the reference's
// operation order (HadrienG2/3photons-rust src/evgen.rs, evcut.rs, spinor.rs,
// matelems.rs, resacc.rs).
#include "tp_kernels.h"
#include "tp_xoshiro.h"

namespace tpk {

using namespace tpd;

struct DevAcc { // layout of tp_acc
    uint64_t selected;
    double spm2[5], vars[5], sigma, variance;
};

// State of (batch b, draw offset off inside the batch): M^off B^b s0.
__device__ __forceinline__ void xo_state_at(const FusedArgs &a, uint64_t b, uint64_t off, uint64_t s[4])
{
    s[0] = a.s0[0];
    s[1] = a.s0[1];
    s[2] = a.s0[2];
    s[3] = a.s0[3];
    for (int k = 0; k < a.kb; ++k)
        if ((b >> k) & 1ull) // uniform across the workgroup
            gf2_apply(a.bt + 1024 * k, s);
    for (int k = 0; k < a.km; ++k)
        if ((off >> k) & 1ull) // per lane
            gf2_apply(a.mt + 1024 * k, s);
}

// ---- per-thread accumulation (resacc.rs:121-129) ---------------------------
// Sums s = (spm2[5], vars[5], sigma, variance); the mixed-precision path
// (PREC 1) keeps Kahan compensations c alongside.
template <int PREC> struct Acc {
    double s[12];
    double c[PREC ? 12 : 1];
    uint32_t n;
};

template <int PREC> __device__ __forceinline__ void acc_zero(Acc<PREC> &A)
{
#pragma unroll
    for (int j = 0; j < 12; ++j)
        A.s[j] = 0.;
#pragma unroll
    for (int j = 0; j < (PREC ? 12 : 1); ++j)
        A.c[j] = 0.;
    A.n = 0;
}

__device__ __forceinline__ void kahan(double &s, double &c, double x)
{
    double y = x - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int PREC> __device__ __forceinline__ void acc_add(Acc<PREC> &A, const Params &P, const double m[5])
{
    double x[12];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        x[k] = m[k];
        x[5 + k] = m[k] * m[k];
    }
    const double w = weight_f64(P, m);
    x[10] = w;
    x[11] = w * w;
    A.n += 1;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        if (PREC == 0)
            A.s[j] = A.s[j] + x[j];
        else
            kahan(A.s[j], A.c[PREC ? j : 0], x[j]);
    }
}

template <int PREC> __device__ __forceinline__ void acc_values(const Acc<PREC> &A, double v[12])
{
#pragma unroll
    for (int j = 0; j < 12; ++j)
        v[j] = PREC ? A.s[j] - A.c[PREC ? j : 0] : A.s[j];
}

// Fixed-shape wavefront butterfly then in-order wave combine; deterministic.
template <int NW>
__device__ __forceinline__ void block_reduce_store(double v[12], uint32_t n, DevAcc *out, double (*red)[13])
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int j = 0; j < 12; ++j)
            v[j] += __shfl_xor(v[j], off);
        n += __shfl_xor(n, off);
    }
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 12; ++j)
            red[wid][j] = v[j];
        red[wid][12] = (double)n;
    }
    __syncthreads();
    if (threadIdx.x < 13) {
        const int j = threadIdx.x;
        double t = red[0][j];
#pragma unroll
        for (int w = 1; w < NW; ++w)
            t = t + red[w][j];
        if (j < 5)
            out->spm2[j] = t;
        else if (j < 10)
            out->vars[j - 5] = t;
        else if (j == 10)
            out->sigma = t;
        else if (j == 11)
            out->variance = t;
        else
            out->selected = (uint64_t)t;
    }
    __syncthreads();
}

// ---- the fused event loop ------------------------------------------------
//
// One wavefront per batch.  Each iteration every lane generates one event
// (uniforms -> RAMBO -> sort slots -> cuts); the photon momenta of the events
// that pass are appended to a queue in LDS (ballot + prefix count), each photon
// into the row of its sort slot, and whenever 64
// are queued the whole wavefront takes them and evaluates spinors, amplitudes
// and weights on full lanes.  Cut events therefore cost no amplitude work, and
// the amplitude code never runs with masked-off lanes except for the final
// partial flush of a batch.  The accumulation order is the queue order, a
// function of the data only, so results are reproducible run to run.
//
// Event assignment inside the batch: with the in-register xoshiro stream each
// lane owns a contiguous chunk (one skip-ahead per chunk); with a stream read
// from HBM lane l takes events l, l+64, ... (coalesced loads).
constexpr int QCAP = 128;

// Three wavefronts per SIMD (163 VGPRs at f64, no scratch): the LDS queue
// (12.4 KB per workgroup) caps a CU at 12 resident workgroups anyway, and the
// third wave hides the transcendental-table, LDS and f64 latencies better than
// two (measured 3 % faster at f64).
// The run constants re-read from the kernel-argument segment (FusedArgs::P
// sits at its start) through a pointer the compiler cannot see through, at
// each use site: scalar loads, which cost no VALU slot, instead of ~40 SGPRs
// held across the whole event loop, which the allocator spills to VGPR lanes
// and reads back with one VALU v_readlane per dword.
static_assert(offsetof(FusedArgs, P) == 0, "Params first in FusedArgs");
typedef const __attribute__((address_space(4))) Params KParams;
__device__ __forceinline__ const Params &fresh_params()
{
    KParams *k = (KParams *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return *(const Params *)k;
}

template <int SOURCE, int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void fused_kernel(FusedArgs a)
{
    __shared__ double queue[12][QCAP];      // momenta, structure of arrays
    __shared__ uint32_t qev[SOURCE == 2 ? QCAP : 1]; // weighted source: the event's index in the chunk
    __shared__ double red[1][13];
    const uint32_t lane = threadIdx.x;
    const uint64_t lanemask_lt = (1ull << lane) - 1ull;
    uint64_t st[4] = {0, 0, 0, 0}; // xoshiro state at this lane's chunk start
    bool have = false;
    DevAcc *out = reinterpret_cast<DevAcc *>(a.out);
    for (uint64_t lb = blockIdx.x; lb < a.nb; lb += gridDim.x) {
        const uint64_t b = a.b_first + lb;
        const uint64_t start = b * a.batch_size;
        const uint64_t left = a.num_events - start;
        const uint32_t n = left < a.batch_size ? (uint32_t)left : a.batch_size;
        const uint32_t q = n / 64, r = n % 64;
        const uint32_t iters = q + (r ? 1u : 0u);
        uint32_t len, e0;
        if (SOURCE == 0) {
            len = q + (lane < r ? 1u : 0u);
            e0 = lane * q + (lane < r ? lane : r);
        } else {
            len = q + (lane < r ? 1u : 0u); // events lane, lane+64, ...
            e0 = lane;
        }
        uint64_t s[4] = {0, 0, 0, 0};
        if (SOURCE == 0) {
            if (!have || n != a.batch_size)
                xo_state_at(a, b, 12ull * e0, st);
            else
                gf2_apply(a.stride, st);
            have = true;
            s[0] = st[0];
            s[1] = st[1];
            s[2] = st[2];
            s[3] = st[3];
        }
        const int32_t *src = (SOURCE == 1) ? a.draws + (size_t)lb * 12 * a.batch_size : nullptr;
        const size_t wbase = (size_t)lb * a.batch_size;
        const RerollBits rb{a.acc1, a.bbase, a.bstride ? (lb + 1) * a.bstride : a.blimit};
        Acc<PREC> A;
        acc_zero(A);
        uint32_t qn = 0; // queued events (uniform)
        // faster-evgen: the lane's next event start, loaded an event ahead
        constexpr bool FE = SOURCE == 3 || SOURCE == 4;
        uint64_t qnext = FE && len > 0 ? a.evq[wbase + lane] : 0;
        for (uint32_t k = 0; k < iters; ++k) {
            bool pass = false;
            double p[3][4];
            int d[3]; // sort slots of the three photons
            const uint64_t qk = qnext;
            if (FE && k + 1 < len)
                qnext = a.evq[wbase + (size_t)((k + 1) * 64 + lane)];
            if (k < len && SOURCE == 3) {
                RanfWalk w{a.draws, a.qbase, qk, rb};
                pass = generate_faster_f64(fresh_params(), w, p, d);
            } else if (k < len && SOURCE == 4) {
                XoWalk w{a.xdraws, a.qbase, qk, rb};
                pass = generate_faster_f64(fresh_params(), w, p, d);
            } else if (k < len) {
                double u[12];
                if (SOURCE == 0) {
#pragma unroll
                    for (int j = 0; j < 12; ++j)
                        u[j] = xo_uniform53(s);
                } else if (SOURCE == 1) {
                    const int32_t *e = src + 12 * (size_t)(k * 64 + lane);
#pragma unroll
                    for (int j = 0; j < 12; ++j)
                        u[j] = (double)e[j] * 1e-9;
                } else {
                    const size_t e = wbase + (size_t)(k * 64 + lane);
#pragma unroll
                    for (int j = 0; j < 12; ++j)
                        u[j] = a.u[j * a.u_stride + e];
                }
                pass = generate_f64<SOURCE == 0>(fresh_params(), u, p, d);
            }
            const uint64_t mask = __ballot(pass);
            if (pass) {
                const uint32_t slot = qn + (uint32_t)__popcll(mask & lanemask_lt);
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        queue[4 * d[i] + j][slot] = p[i][j];
                if (SOURCE == 2)
                    qev[slot] = (uint32_t)(wbase + (size_t)(k * 64 + lane));
            }
            qn += (uint32_t)__popcll(mask);
            if (qn >= 64) {
                __syncthreads();
                const uint32_t slot = qn - 64 + lane;
                double pq[3][4], m[5];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        pq[i][j] = queue[4 * i + j][slot];
                const double wq = SOURCE == 2 ? a.invq[qev[slot]] : 1.;
                __syncthreads();
                qn -= 64;
                matrix_elements<PREC>(fresh_params(), pq, m);
                if (SOURCE == 2)
#pragma unroll
                    for (int c = 0; c < 5; ++c)
                        m[c] = m[c] * wq;
                acc_add(A, fresh_params(), m);
            }
        }
        if (qn > 0) { // partial flush
            __syncthreads();
            if (lane < qn) {
                double pq[3][4], m[5];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        pq[i][j] = queue[4 * i + j][lane];
                matrix_elements<PREC>(fresh_params(), pq, m);
                if (SOURCE == 2) {
                    const double wq = a.invq[qev[lane]];
#pragma unroll
                    for (int c = 0; c < 5; ++c)
                        m[c] = m[c] * wq;
                }
                acc_add(A, fresh_params(), m);
            }
        }
        __syncthreads();
        double v[12];
        acc_values(A, v);
        block_reduce_store<1>(v, A.n, out + lb, red);
    }
}

template <int SOURCE, int PREC> int blocks_per_cu()
{
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_kernel<SOURCE, PREC>, 64, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

int fused_blocks_per_cu(int source, int precision)
{
    if (source == 0)
        return precision == 0 ? blocks_per_cu<0, 0>() : blocks_per_cu<0, 1>();
    if (source == 2)
        return blocks_per_cu<2, 0>();
    if (source == 3)
        return blocks_per_cu<3, 0>();
    if (source == 4)
        return blocks_per_cu<4, 0>();
    return precision == 0 ? blocks_per_cu<1, 0>() : blocks_per_cu<1, 1>();
}

void launch_fused(const FusedArgs &a, int source, int grid, hipStream_t st)
{
    const dim3 g(grid), b(64);
    if (source == 0) {
        if (a.precision == 0)
            hipLaunchKernelGGL((fused_kernel<0, 0>), g, b, 0, st, a);
        else
            hipLaunchKernelGGL((fused_kernel<0, 1>), g, b, 0, st, a);
    } else if (source == 1) {
        if (a.precision == 0)
            hipLaunchKernelGGL((fused_kernel<1, 0>), g, b, 0, st, a);
        else
            hipLaunchKernelGGL((fused_kernel<1, 1>), g, b, 0, st, a);
    } else if (source == 2) {
        hipLaunchKernelGGL((fused_kernel<2, 0>), g, b, 0, st, a);
    } else if (source == 3) {
        hipLaunchKernelGGL((fused_kernel<3, 0>), g, b, 0, st, a);
    } else {
        hipLaunchKernelGGL((fused_kernel<4, 0>), g, b, 0, st, a);
    }
}

template <int SOURCE, int PREC>
__global__ __launch_bounds__(256) void events_kernel(FusedArgs a, uint32_t n, EventOut o)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    double u[12];
    if (SOURCE == 0) {
        uint64_t s[4];
        s[0] = a.s0[0];
        s[1] = a.s0[1];
        s[2] = a.s0[2];
        s[3] = a.s0[3];
        const uint64_t off = 12ull * e;
        for (int k = 0; k < a.km; ++k)
            if ((off >> k) & 1ull)
                gf2_apply(a.mt + 1024 * k, s);
#pragma unroll
        for (int j = 0; j < 12; ++j)
            u[j] = xo_uniform53(s); // u 2^53, as the fused kernel draws them
    } else {
#pragma unroll
        for (int j = 0; j < 12; ++j)
            u[j] = (double)a.draws[12ull * e + j] * 1e-9;
    }
    double pu[3][4], p[3][4], m[5] = {0., 0., 0., 0., 0.};
    int d[3];
    const bool k = generate_f64<SOURCE == 0>(a.P, u, pu, d);
    to_slots(pu, d, p);
    if (SOURCE == 0)
#pragma unroll
        for (int j = 0; j < 12; ++j)
            u[j] = u[j] * 0x1.0p-53;
    if (k)
        matrix_elements<PREC>(a.P, p, m);
    if (!k) {
#pragma unroll
        for (int c = 0; c < 5; ++c)
            m[c] = 0.;
    }
    if (o.u)
#pragma unroll
        for (int j = 0; j < 12; ++j)
            o.u[12ull * e + j] = u[j];
    if (o.p)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                o.p[12ull * e + 4 * i + j] = p[i][j];
    if (o.keep)
        o.keep[e] = k ? 1 : 0;
    if (o.m)
#pragma unroll
        for (int c = 0; c < 5; ++c)
            o.m[5ull * e + c] = m[c];
}

void launch_events(const FusedArgs &a, int source, uint32_t n, const EventOut &o, hipStream_t st)
{
    const dim3 g((n + 255) / 256), b(256);
    if (source == 0) {
        if (a.precision == 0)
            hipLaunchKernelGGL((events_kernel<0, 0>), g, b, 0, st, a, n, o);
        else
            hipLaunchKernelGGL((events_kernel<0, 1>), g, b, 0, st, a, n, o);
    } else {
        if (a.precision == 0)
            hipLaunchKernelGGL((events_kernel<1, 0>), g, b, 0, st, a, n, o);
        else
            hipLaunchKernelGGL((events_kernel<1, 1>), g, b, 0, st, a, n, o);
    }
}

// ---- RANF expansion ------------------------------------------------------
constexpr int32_t MODULO = 1000000000;

__device__ __forceinline__ int32_t norm_mod(int32_t v)
{
    v += (v < 0) ? MODULO : 0;
    v += (v < 0) ? MODULO : 0;
    v -= (v >= MODULO) ? MODULO : 0;
    v -= (v >= MODULO) ? MODULO : 0;
    return v;
}

// One RANF round in closed form (lane i - 1 holds X[i], i = 1..55):
//   i <= 24:  X'[i] = X[i] - X[i+31]
//   25..48:   X'[i] = X[i] - X'[i-24] = X[i] - X[i-24] + X[i+7]
//   49..55:   X'[i] = X[i] - X[i-24] + X[i-48] - X[i-17]
// (ranf.rs:106-119 applies the first two lines as in-place updates; the third
// is the second with X'[i-24] expanded once more), all mod 1e9.  Every lane
// evaluates (X[i] - b) + (c - d) with b, c, d read from other lanes of the
// same wave by ds_bpermute (no LDS storage, no barrier).  Lanes 55..63 hold 0
// and keep it, so a term a case does not have reads lane 63; each value is the
// unique representative in [0, 1e9), so the sum order does not matter.  The
// whole wave must be active: one wave holds one window.
__device__ __forceinline__ uint32_t mod_sub(uint32_t a, uint32_t b) // (a - b) mod 1e9, a, b in [0, 1e9)
{
    const uint32_t t = a - b;
    return min(t, t + (uint32_t)MODULO);
}
struct RanfLanes { // byte addresses for ds_bpermute of the lanes a lane reads
    int b, c, d;
};
__device__ __forceinline__ RanfLanes ranf_lanes()
{
    const int lane = (int)(threadIdx.x & 63u), i = lane + 1;
    const bool live = lane < 55;
    RanfLanes l;
    l.b = (live ? (i <= 24 ? i + 30 : i - 25) : 63) << 2;          // X[i+31] or X[i-24]
    l.c = (live && i > 24 ? (i <= 48 ? i + 6 : i - 49) : 63) << 2; // X[i+7] or X[i-48]
    l.d = (live && i > 48 ? i - 18 : 63) << 2;                     // X[i-17]
    return l;
}
__device__ __forceinline__ int32_t ranf_next_round(int32_t x, const RanfLanes &l)
{
    const uint32_t wb = (uint32_t)__builtin_amdgcn_ds_bpermute(l.b, x);
    const uint32_t wc = (uint32_t)__builtin_amdgcn_ds_bpermute(l.c, x);
    const uint32_t wd = (uint32_t)__builtin_amdgcn_ds_bpermute(l.d, x);
    const uint32_t v = mod_sub((uint32_t)x, wb) + mod_sub(wc, wd); // [0, 2e9)
    return (int32_t)min(v, v - (uint32_t)MODULO);
}

// X_{r+m} from X_r with c = x^(55 m) mod P: X_{r+m}[i] = sum_k c_k Y[i-1+k]
// over the window Y = (X_r, X_{r+1}).
__device__ __forceinline__ int32_t ranf_jump(int32_t x, const uint32_t *__restrict__ c, int32_t *win)
{
    const int lane = threadIdx.x;
    int32_t nx = ranf_next_round(x, ranf_lanes());
    if (lane < 55) {
        win[lane] = x;
        win[55 + lane] = nx;
    }
    __syncthreads();
    uint64_t acc = 0;
    if (lane < 55) {
#pragma unroll 5
        for (int k = 0; k < 55; ++k) {
            acc += (uint64_t)c[k] * (uint64_t)(uint32_t)win[lane + k];
            if ((k & 15) == 15)
                acc %= (uint64_t)MODULO;
        }
        acc %= (uint64_t)MODULO;
    }
    __syncthreads();
    return lane < 55 ? (int32_t)acc : 0;
}

// Outcome of the unit-disc test on the point (x, y) = 2u - 1 of two
// uniforms: r2 above 1 re-rolls, and in f64 so does r2 below MIN_POSITIVE^2
// (f32's MIN_POSITIVE^2 is 0).  The walk sums (0 + x*x) + y*y for the six
// values' points and 0 + (x*x + y*y) for a re-roll: without contraction both
// are fl(fl(x*x) + fl(y*y)).
__device__ __forceinline__ uint32_t disc_accept(double x, double y)
{
    const double min2 = 2.2250738585072014e-308 * 2.2250738585072014e-308;
    const double r2 = (0. + x * x) + y * y;
    return (r2 > 1. || r2 < min2) ? 0u : 1u;
}
__device__ __forceinline__ uint32_t disc_accept(float x, float y)
{
    const float r2 = (0.f + x * x) + y * y;
    return r2 > 1.f ? 0u : 1u;
}

// A RANF value as the walk's unit-disc test sees it: x = 2 (v * 1e-9) - 1 in
// the run's float type (ranf.rs:99), and its square.
template <class T> __device__ __forceinline__ T ranf_sq(int32_t v)
{
    const T x = fma((T)v * (T)1e-9, (T)2., (T)-1.); // 2u is exact: one rounding either way
    return x * x;
}

// Lane moves within the wavefront (DPP, no LDS): DPP_WAVE_ROR1 gives lane i
// the value of lane i - 1 (lane 0: lane 63); DPP_ROW_SHR9 gives lane i of each
// row of 16 the value of lane i - 9 of the row (lanes 61..63: lanes 52..54).
constexpr int DPP_WAVE_ROR1 = 0x13C, DPP_ROW_SHR9 = 0x119;
template <int CTRL> __device__ __forceinline__ double lanes_dpp(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL> __device__ __forceinline__ float lanes_dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Accept bits of a segment's rounds.  Lane i holds the square of position
// 55 r + 54 - i (each value's square is computed once, as the next round's,
// and carried); its partners +1 and +3 come from lanes i - 1 and i - 3, or
// for the round's last positions from the next round's first ones, by lane
// moves.  Each round's two ballots, bit-reversed into position order, wait in
// LDS; every 32 rounds (1760 positions, 55 words) lanes 0..54 assemble one
// word each and store them together (barriers order lane 0's writes and the
// other lanes' reads: the workgroup is the one wavefront).
template <class T> struct RanfBits {
    uint64_t *m1, *m3;  // [32] LDS: the block's rounds' bits, bit j = position 55 r + j
    uint32_t *d1, *d3;  // the block's first words
    uint32_t k;         // rounds of the block so far
    T sq;               // this round's squares
    __device__ __forceinline__ void start(int32_t x) { sq = ranf_sq<T>(x); }
    __device__ __forceinline__ void round(int32_t xn)
    {
        const int lane = (int)(threadIdx.x & 63u);
        const T sqn = ranf_sq<T>(xn);
        // t: this round's squares in lanes 0..54 and the next round's
        // positions 2, 1, 0 (its lanes 52..54) in lanes 61..63; rotating t
        // down the wavefront by 1 and by 3 lanes gives each lane its partners
        // +1 and +3
        // (the empty asm keeps the move out of the select's branch: every
        // lane takes part, so the source lanes 52..54 are live)
        T nx = lanes_dpp<DPP_ROW_SHR9>(sqn);
        asm volatile("" : "+v"(nx));
        const T t = lane < 55 ? sq : nx;
        const T p1 = lanes_dpp<DPP_WAVE_ROR1>(t);
        const T p3 = lanes_dpp<DPP_WAVE_ROR1>(lanes_dpp<DPP_WAVE_ROR1>(p1));
        // r2 = fl(fl(x*x) + fl(y*y)), the walk's sums (0 + x*x) + y*y and
        // 0 + (x*x + y*y); a point re-rolls when r2 > 1 (the f64 test's
        // r2 < MIN_POSITIVE^2 compares with 0, which a sum of squares never
        // is below)
        const T r1 = sq + p1, r3 = sq + p3;
        constexpr uint64_t LIVE = (1ull << 55) - 1;
        const uint64_t b1 = __builtin_bitreverse64(__ballot(r1 <= (T)1.) & LIVE) >> 9;
        const uint64_t b3 = __builtin_bitreverse64(__ballot(r3 <= (T)1.) & LIVE) >> 9;
        if (lane == 0) {
            m1[k] = b1;
            m3[k] = b3;
        }
        sq = sqn;
        if (++k == 32)
            flush();
    }
    // words j of the block with 32 j < 55 k (those past the segment's last
    // round hold positions no walk reads)
    __device__ __forceinline__ void flush()
    {
        const uint32_t lane = threadIdx.x & 63u;
        __syncthreads(); // lane 0's masks are in
        if (lane < 55 && 32 * lane < 55 * k) {
            const uint32_t r0 = 32 * lane / 55, b0 = 32 * lane % 55;
            uint64_t w1 = m1[r0] >> b0, w3 = m3[r0] >> b0;
            if (55 - b0 < 32) {
                w1 |= m1[r0 + 1] << (55 - b0);
                w3 |= m3[r0 + 1] << (55 - b0);
            }
            d1[lane] = (uint32_t)w1;
            d3[lane] = (uint32_t)w3;
        }
        d1 += 55;
        d3 += 55;
        k = 0;
        __syncthreads(); // before the next block's masks overwrite these
    }
};

template <int BITS> __global__ __launch_bounds__(64) void ranf_expand_kernel(RanfArgs a)
{
    __shared__ int32_t win[112];
    const int lane = threadIdx.x;
    const uint64_t piece = blockIdx.x / a.nsegs;
    const uint64_t seg = a.seg_first + blockIdx.x % a.nsegs;
    const uint64_t p0 = (uint64_t)__builtin_amdgcn_readfirstlane(a.p0[piece]); // uniform: scalar compares below
    const uint64_t q_begin = p0 + a.d_begin, q_end = p0 + a.d_end;
    const uint64_t r_begin = seg * (uint64_t)a.R;
    if (r_begin * 55 >= q_end)
        return; // whole workgroup
    int32_t x = lane < 55 ? a.origin[56 * piece + lane + 1] : 0;
    for (int k = 0; k < a.K; ++k)
        if ((seg >> k) & 1ull)
            x = ranf_jump(x, a.polys + 55 * k, win);
    int32_t *out = a.out + piece * a.piece_stride;
    const RanfLanes l = ranf_lanes();
    __shared__ uint64_t bm[2][BITS ? 32 : 1];
    using BT = typename std::conditional<BITS == 2, float, double>::type;
    RanfBits<BT> bits;
    if (BITS) {
        const uint64_t w0 = piece * a.bits_piece_words + (seg - a.seg_first) * (55ull * (uint64_t)a.R / 32);
        bits.m1 = bm[0];
        bits.m3 = bm[1];
        bits.d1 = a.acc1 + w0;
        bits.d3 = a.acc3 + w0;
        bits.k = 0;
        bits.start(x);
    }
    // the round's accept bits (from it and the next round), then on to the next
    auto advance = [&](int32_t x) -> int32_t {
        const int32_t xn = ranf_next_round(x, l);
        if (BITS)
            bits.round(xn);
        return xn;
    };
    // X[lane+1] of round r is emitted at stream position 55 r + 54 - lane.
    // (A round's store is 220 bytes, not whole lines: the L2 merges them, so
    // they are ordinary stores; streaming stores took twice as long.)
    // Rounds [r_begin, r_end) reach into [q_begin, q_end); of those,
    // [f_lo, f_hi) lie wholly inside and are stored without range checks.
    const uint64_t r_end = min(r_begin + (uint64_t)a.R, (q_end + 54) / 55);
    const uint64_t f_lo = min(max((q_begin + 54) / 55, r_begin), r_end);
    const uint64_t f_hi = max(f_lo, min(r_end, q_end / 55));
    const uint64_t qq0 = (uint64_t)(54 - lane);
    uint64_t r = r_begin;
    for (; r < f_lo; ++r) {
        const uint64_t qq = 55 * r + qq0;
        if (lane < 55 && qq >= q_begin && qq < q_end)
            out[qq - q_begin] = x;
        x = advance(x);
    }
    int32_t *row = out + (55 * f_lo - q_begin); // uniform
    const int off = 54 - lane;
    const uint32_t nfull = (uint32_t)(f_hi - f_lo); // at most R
    for (uint32_t k = 0; k < nfull; ++k) {
        if (lane < 55)
            row[off] = x;
        x = advance(x);
        row += 55;
    }
    r = f_hi;
    for (; r < r_end; ++r) {
        const uint64_t qq = 55 * r + qq0;
        if (lane < 55 && qq >= q_begin && qq < q_end)
            out[qq - q_begin] = x;
        x = advance(x);
    }
    if (BITS && bits.k)
        bits.flush();
}

void launch_ranf_expand(const RanfArgs &a, uint64_t npieces, hipStream_t st)
{
    const uint64_t blocks = npieces * a.nsegs;
    if (a.bits == 1)
        hipLaunchKernelGGL(ranf_expand_kernel<1>, dim3((unsigned)blocks), dim3(64), 0, st, a);
    else if (a.bits == 2)
        hipLaunchKernelGGL(ranf_expand_kernel<2>, dim3((unsigned)blocks), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL(ranf_expand_kernel<0>, dim3((unsigned)blocks), dim3(64), 0, st, a);
}

// ---- faster-evgen stream walk --------------------------------------------------
// take(n) of random_array::<N> on the stored stream: a request that does not
// fit in the current round skips its leftovers (ranf.rs:89-92).  False (and no
// read) when the values would lie past the stored range.  T is the run's
// float type: (n as T) * 1e-9 (ranf.rs:99).
template <class T> __device__ __forceinline__ bool fe_take(const FeWalkArgs &a, uint64_t &q, int n, T *out)
{
    const uint32_t rem = 55u - (uint32_t)(q % 55u);
    if (rem < (uint32_t)n)
        q += rem;
    if (q + (uint64_t)n > a.qlimit)
        return false;
    if (out) {
        const int32_t *v = a.buf + (q - a.qbase);
        for (int k = 0; k < n; ++k)
            out[k] = (T)v[n - 1 - k] * (T)1e-9;
    }
    q += (uint64_t)n;
    return true;
}

// Consumption of one event (raw_photons_faster_f64's draws: 9, 6, then 2 per
// re-rolled point of the unit-disc test, decided in the run's float type T).
// False when the stored range ends inside the event.
template <class T = double> __device__ __forceinline__ bool fe_walk_event(const FeWalkArgs &a, uint64_t &q)
{
    T r6[6];
    if (!fe_take<T>(a, q, 9, nullptr) || !fe_take<T>(a, q, 6, r6))
        return false;
    const T minpos = sizeof(T) == 8 ? (T)2.2250738585072014e-308 : (T)1.17549435e-38f;
    const T min2 = minpos * minpos; // 0 in f32, as in the reference's f32 build
    for (int i = 0; i < 3; ++i) {
        const T x = (T)2. * r6[i] - (T)1., y = (T)2. * r6[3 + i] - (T)1.;
        T r2 = ((T)0. + x * x) + y * y;
        while (r2 > (T)1. || r2 < min2) {
            T r[2];
            if (!fe_take<T>(a, q, 2, r))
                return false;
            const T xa = (T)2. * r[0] - (T)1., ya = (T)2. * r[1] - (T)1.;
            r2 = (T)0. + (xa * xa + ya * ya);
        }
    }
    return true;
}

// Accept bit of stored position r (bit r % 32 of word r / 32).
__device__ __forceinline__ uint32_t acc_bit(const uint32_t *bits, uint64_t r) { return (bits[r >> 5] >> (r & 31)) & 1u; }

// Of the accepted re-roll pairs marked in y (bit k: the pair starting k
// positions on), the fails-th one (fails in 1..3): clearing the lowest
// fails - 1 set bits leaves it as the lowest.  0 when y has fewer.
__device__ __forceinline__ uint64_t nth_accept(uint64_t y, uint32_t fails)
{
    const uint64_t y1 = y & (y - 1), y2 = y1 & (y1 - 1);
    return fails == 1 ? y : fails == 2 ? y1 : y2;
}

// 64 bits of a bit array from bit r on: the words at w[0], w[1], w[2] hold
// bits from (r & ~31) on.
__device__ __forceinline__ uint64_t bits64(const uint32_t *w, uint64_t r)
{
    const uint32_t s = (uint32_t)(r & 31);
    const uint64_t lo = ((uint64_t)w[1] << 32) | w[0];
    return s ? (lo >> s) | ((uint64_t)w[2] << (64 - s)) : lo;
}

// One event's consumption of a xoshiro stream (xoshiro256+ in f64 or
// xoshiro128+ in f32): 9, 6, then 2 per re-rolled point of the unit-disc
// test, no discards.  The test's outcomes come from the accept bits (bit
// index = position - bbase): point i of the 6 values is (u_p, u_p+3) at
// p = q + 9 + i (acc3), a re-roll is (u_p, u_p+1) (acc1).  The re-rolls of
// the failed points form one run of pairs from q + 15 on, so with F failed
// points the event ends after the F-th accepted pair, found from 64 positions
// of acc1 bits at once (the pairs are the even offsets).  Near qlimit, or when
// the 32 pairs hold fewer than F accepted ones, the pairs are taken one at a
// time.  Every bit read pairs positions below qlimit.  False when the stored
// range ends inside the event.
__device__ __forceinline__ bool xo_walk_event(const uint32_t *acc1, const uint32_t *acc3, uint64_t bbase,
                                              uint64_t qlimit, uint64_t &q)
{
    if (q + 111 <= qlimit) { // words up to position q + 15 + 95 lie inside the stored range
        const uint64_t r3 = q + 9 - bbase, r1 = q + 15 - bbase;
        const uint32_t *p3 = acc3 + (r3 >> 5);
        const uint32_t ok3 = (uint32_t)((((uint64_t)p3[1] << 32) | p3[0]) >> (r3 & 31)) & 7u;
        const uint32_t fails = 3u - (uint32_t)__popc(ok3);
        if (fails == 0) {
            q += 15;
            return true;
        }
        const uint64_t y = nth_accept(bits64(acc1 + (r1 >> 5), r1) & 0x5555555555555555ull, fails);
        if (y) {
            q += 17 + (uint64_t)__builtin_ctzll(y);
            return true;
        }
    }
    q += 9;
    if (q + 6 > qlimit)
        return false;
    const uint64_t r = q - bbase;
    q += 6;
    for (int i = 0; i < 3; ++i) {
        if (acc_bit(acc3, r + i))
            continue;
        for (;;) {
            if (q + 2 > qlimit)
                return false;
            const uint32_t ok = acc_bit(acc1, q - bbase);
            q += 2;
            if (ok)
                break;
        }
    }
    return true;
}

// The RANF stream (frame positions: place in the round a = q % 55): take(n)
// first skips the round's leftovers when fewer than n remain (ranf.rs:89-92).
__device__ __forceinline__ void ranf_skip(uint64_t &q, uint32_t &a, uint32_t n)
{
    const uint32_t rem = 55u - a;
    if (rem < n) {
        q += rem;
        a = 0;
    }
}
__device__ __forceinline__ void ranf_step(uint64_t &q, uint32_t &a, uint32_t n)
{
    q += n;
    a += n;
    if (a == 55u)
        a = 0;
}

// One event's consumption of the RANF stream on its accept bits, as
// fe_walk_event: take(9), take(6) (its points are (p, p+3) for p the block's
// first three positions, in any order: only the count of failures matters),
// then a take(2) per re-roll until each failed point has an accepted pair.
__device__ __forceinline__ bool ranf_walk_event(const uint32_t *acc1, const uint32_t *acc3, uint64_t bbase,
                                                uint64_t qlimit, uint64_t &q, uint32_t &a)
{
    ranf_skip(q, a, 9);
    if (q + 9 > qlimit)
        return false;
    ranf_step(q, a, 9);
    ranf_skip(q, a, 6);
    if (q + 6 > qlimit)
        return false;
    const uint64_t r = q - bbase;
    ranf_step(q, a, 6);
    uint32_t fails = 3u - (acc_bit(acc3, r) + acc_bit(acc3, r + 1) + acc_bit(acc3, r + 2));
    while (fails) {
        ranf_skip(q, a, 2);
        if (q + 2 > qlimit)
            return false;
        fails -= acc_bit(acc1, q - bbase);
        ranf_step(q, a, 2);
    }
    return true;
}

// The walkers' accept bits through LDS.  Bits read straight from HBM cost
// each event several scattered loads, and a compute unit's walkers touch
// more lines than its L1 holds.  So each walking lane keeps BW_WORDS words of
// acc1 and acc3 (BW_WORDS * 32 positions) from the word of its next event's
// first test on in rows of its own, and takes its events' tests from there;
// the wavefront refills every lane's rows (16-byte loads) when some lane's
// next event could leave its window.  No lane reads another's rows.
constexpr int BW_WORDS = 32, BW_ROW = BW_WORDS + 4, BW_THREADS = 128;

struct BitWin {
    uint32_t *w1, *w3; // the lane's rows
    uint64_t wb;       // first word held (counted from bbase)
    uint64_t end;      // stream position where the held bits end
};

// Words from the one holding position q + 9 on, in whole groups of 4 below
// qlimit.
__device__ __forceinline__ void bit_win_fill(BitWin &v, const uint32_t *acc1, const uint32_t *acc3, uint64_t bbase,
                                             uint64_t qlimit, uint64_t q)
{
    const uint64_t lim4 = ((qlimit - bbase) >> 5) & ~3ull;
    const uint64_t wb = ((q + 9 - bbase) >> 5) & ~3ull;
    uint64_t have = 0;
#pragma unroll
    for (int i = 0; i < BW_WORDS / 4; ++i) {
        const uint64_t w = wb + 4 * i;
        if (w < lim4) {
            *reinterpret_cast<uint4 *>(v.w1 + 4 * i) = *reinterpret_cast<const uint4 *>(acc1 + w);
            *reinterpret_cast<uint4 *>(v.w3 + 4 * i) = *reinterpret_cast<const uint4 *>(acc3 + w);
            have = 4 * i + 4;
        }
    }
    v.wb = wb;
    v.end = bbase + 32 * (wb + have);
}

// xo_walk_event with the tests from the lane's window when the event's bits
// (as far as position q + 110) lie in it.
struct XoBitWalk {
    static constexpr uint64_t REACH = 111;
    const uint32_t *acc1, *acc3;
    uint64_t bbase, qlimit;
    __device__ __forceinline__ void start(uint64_t) {}
    __device__ __forceinline__ bool event(const BitWin &v, uint64_t &q)
    {
        if (q + REACH <= v.end) {
            const uint64_t r3 = q + 9 - bbase, r1 = q + 15 - bbase;
            const uint32_t *p3 = v.w3 + ((r3 >> 5) - v.wb);
            const uint32_t ok3 = (uint32_t)((((uint64_t)p3[1] << 32) | p3[0]) >> (r3 & 31)) & 7u;
            const uint32_t fails = 3u - (uint32_t)__popc(ok3);
            if (fails == 0) {
                q += 15;
                return true;
            }
            const uint64_t y = nth_accept(bits64(v.w1 + ((r1 >> 5) - v.wb), r1) & 0x5555555555555555ull, fails);
            if (y) {
                q += 17 + (uint64_t)__builtin_ctzll(y);
                return true;
            }
        }
        return xo_walk_event(acc1, acc3, bbase, qlimit, q);
    }
};

// ranf_walk_event with the tests from the lane's window when the event's
// bits (as far as position q + 123: up to 13 discarded leftovers, the 15
// values, 64 positions of re-rolls and the rest of their last word) lie in
// it: the 9 and 6 takes by position arithmetic, the re-rolls as the F-th
// accepted pair among ranf_pair_starts.
struct RanfBitWalk {
    static constexpr uint64_t REACH = 128;
    const uint32_t *acc1, *acc3;
    uint64_t bbase, qlimit;
    uint32_t a; // q % 55
    __device__ __forceinline__ void start(uint64_t q) { a = (uint32_t)(q % 55u); }
    __device__ __forceinline__ bool event(const BitWin &v, uint64_t &q)
    {
        if (q + REACH <= v.end) {
            uint64_t q1 = q;
            uint32_t a1 = a;
            ranf_skip(q1, a1, 9);
            ranf_step(q1, a1, 9);
            ranf_skip(q1, a1, 6);
            const uint64_t r3 = q1 - bbase;
            ranf_step(q1, a1, 6);
            const uint32_t *p3 = v.w3 + ((r3 >> 5) - v.wb);
            const uint32_t ok3 = (uint32_t)((((uint64_t)p3[1] << 32) | p3[0]) >> (r3 & 31)) & 7u;
            const uint32_t fails = 3u - (uint32_t)__popc(ok3);
            if (fails == 0) {
                q = q1;
                a = a1;
                return true;
            }
            const uint64_t r1 = q1 - bbase;
            const uint64_t y = nth_accept(bits64(v.w1 + ((r1 >> 5) - v.wb), r1) & ranf_pair_starts(55u - a1), fails);
            if (y) {
                const uint32_t k = (uint32_t)__builtin_ctzll(y) + 2; // past the accepted pair
                q = q1 + k;
                a = a1 + k;
                a = a >= 110u ? a - 110u : a >= 55u ? a - 55u : a;
                return true;
            }
        }
        return ranf_walk_event(acc1, acc3, bbase, qlimit, q, a);
    }
};

// Records starts from q while q < stop and the list has room (each event
// consumes at least 15 positions, so the loop ends); returns the count.
template <class W>
__device__ __forceinline__ uint32_t bit_walk_run(W w, BitWin &v, uint64_t q, uint64_t stop, uint64_t *list,
                                                 uint32_t cap)
{
    v.wb = 0;
    v.end = 0;
    w.start(q);
    uint32_t n = 0;
    while (q < stop && n < cap) {
        list[n++] = q;
        if (__any(q + W::REACH > v.end && q + W::REACH <= w.qlimit))
            bit_win_fill(v, w.acc1, w.acc3, w.bbase, w.qlimit, q);
        if (!w.event(v, q))
            break;
    }
    return n;
}

// One event of the walk of kind a.kind: RANF with the unit-disc test in f64
// (0) or f32 (1), the xoshiro256+ stream (2) or the xoshiro128+ one (3).
__device__ __forceinline__ bool fe_walk_any(const FeWalkArgs &a, uint64_t &q)
{
    if (a.kind >= 2)
        return xo_walk_event(a.acc1, a.acc3, a.bbase, a.qlimit, q);
    return a.kind == 1 ? fe_walk_event<float>(a, q) : fe_walk_event<double>(a, q);
}

// One thread per walker.  Every step records a start and then consumes at
// least 15 positions, and the loop stops at the segment end, the list
// capacity or the stored range, so it always ends.  The tests come from the
// accept bits through LDS windows (every kind).
__global__ __launch_bounds__(BW_THREADS) void fe_bit_walk_kernel(FeWalkArgs a)
{
    __shared__ uint32_t win[2 * BW_THREADS * BW_ROW];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nwalkers)
        return;
    BitWin v;
    v.w1 = win + threadIdx.x * BW_ROW;
    v.w3 = win + (BW_THREADS + threadIdx.x) * BW_ROW;
    const uint64_t q = k == 0 ? a.q_start : a.seg_origin + (uint64_t)k * a.seg;
    const uint64_t stop = a.seg_origin + (uint64_t)(k + 1) * a.seg + a.overlap;
    uint64_t *list = a.lists + (size_t)k * a.cap;
    a.counts[k] = a.kind >= 2 ? bit_walk_run(XoBitWalk{a.acc1, a.acc3, a.bbase, a.qlimit}, v, q, stop, list, a.cap)
                              : bit_walk_run(RanfBitWalk{a.acc1, a.acc3, a.bbase, a.qlimit, 0}, v, q, stop, list, a.cap);
}

// Walker k's first start that walker k-1 also reached past segment k's start:
// from there on both walks agree.
__global__ __launch_bounds__(256) void fe_join_kernel(FeWalkArgs a)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nwalkers)
        return;
    if (k == 0) {
        a.join_prev[0] = a.join_next[0] = 0;
        return;
    }
    const uint64_t *pa = a.lists + (size_t)(k - 1) * a.cap, *pb = a.lists + (size_t)k * a.cap;
    const uint32_t na = a.counts[k - 1], nb = a.counts[k];
    const uint64_t qs = a.seg_origin + (uint64_t)k * a.seg;
    uint32_t lo = 0, hi = na; // first index of list k-1 with a start >= qs
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (pa[mid] < qs)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t i = lo, j = 0;
    int32_t ji = -1, jj = -1;
    while (i < na && j < nb) {
        if (pa[i] == pb[j]) {
            ji = (int32_t)i;
            jj = (int32_t)j;
            break;
        }
        if (pa[i] < pb[j])
            ++i;
        else
            ++j;
    }
    a.join_prev[k] = ji;
    a.join_next[k] = jj;
}

void launch_fe_walk(const FeWalkArgs &a, hipStream_t st)
{
    const unsigned blocks = (a.nwalkers + 255) / 256;
    hipLaunchKernelGGL(fe_bit_walk_kernel, dim3((a.nwalkers + BW_THREADS - 1) / BW_THREADS), dim3(BW_THREADS), 0, st,
                       a);
    hipLaunchKernelGGL(fe_join_kernel, dim3(blocks), dim3(256), 0, st, a);
}

// Walker k's piece of the true walk is its list from where walker k - 1 met
// it (walker 0: from its start) to where it meets walker k + 1 (the last
// walker: to its end); it must meet walker k + 1 after that point.  The
// pieces are cut to the first `need` starts: off = min(prefix, need).  Three
// passes over blocks of FE_STITCH_BLOCK walkers (at most FE_STITCH_THREADS
// blocks): each piece and its block's sum and failure; the scan of the block
// sums (one workgroup) and the status; each block's own scan, the cut and
// start need - 1.  status[2 ..] holds the blocks' sums (then offsets) and
// status[2 + FE_STITCH_THREADS ..] their failures.
constexpr int FE_STITCH_BLOCK = 256;

__global__ __launch_bounds__(FE_STITCH_BLOCK) void fe_stitch_pieces_kernel(FeWalkArgs a, uint32_t *beg, uint32_t *len,
                                                                            uint64_t *status)
{
    __shared__ uint64_t s[FE_STITCH_BLOCK];
    __shared__ uint32_t anybad;
    const uint32_t T = a.nwalkers, k = blockIdx.x * FE_STITCH_BLOCK + threadIdx.x;
    if (threadIdx.x == 0)
        anybad = 0;
    uint32_t l = 0;
    bool ok = true;
    if (k < T) {
        const uint32_t b = k == 0 ? 0u : (uint32_t)a.join_next[k];
        const int64_t e = k + 1 < T ? (int64_t)a.join_prev[k + 1] : (int64_t)a.counts[T - 1];
        ok = e >= 0 && (uint64_t)e >= (uint64_t)b;
        l = ok ? (uint32_t)(e - (int64_t)b) : 0u;
        beg[k] = b;
        len[k] = l;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    if (!ok)
        atomicOr(&anybad, 1u);
    for (uint32_t d = FE_STITCH_BLOCK / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d)
            s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        status[2 + blockIdx.x] = s[0];
        status[2 + FE_STITCH_THREADS + blockIdx.x] = anybad;
    }
}

__global__ __launch_bounds__(FE_STITCH_THREADS) void fe_stitch_scan_kernel(uint32_t nblocks, uint64_t need,
                                                                           uint64_t *status)
{
    __shared__ uint64_t part[FE_STITCH_THREADS];
    __shared__ uint32_t anybad;
    const uint32_t t = threadIdx.x;
    if (t == 0)
        anybad = 0;
    const uint64_t v = t < nblocks ? status[2 + t] : 0;
    part[t] = v;
    __syncthreads();
    if (t < nblocks && status[2 + FE_STITCH_THREADS + t])
        atomicOr(&anybad, 1u);
    for (uint32_t d = 1; d < FE_STITCH_THREADS; d <<= 1) { // inclusive scan
        const uint64_t u = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    if (t < nblocks)
        status[2 + t] = part[t] - v; // the block's first walker's prefix
    if (t == 0)
        status[0] = (anybad || part[FE_STITCH_THREADS - 1] < need) ? 1u : 0u;
}

__global__ __launch_bounds__(FE_STITCH_BLOCK) void fe_stitch_cut_kernel(FeWalkArgs a, uint64_t need, const uint32_t *beg,
                                                                         uint32_t *len, uint64_t *off, uint64_t *status)
{
    __shared__ uint64_t s[FE_STITCH_BLOCK];
    const uint32_t T = a.nwalkers, t = threadIdx.x, k = blockIdx.x * FE_STITCH_BLOCK + t;
    const uint32_t l = k < T ? len[k] : 0u;
    s[t] = l;
    __syncthreads();
    for (uint32_t d = 1; d < FE_STITCH_BLOCK; d <<= 1) { // inclusive scan
        const uint64_t u = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += u;
        __syncthreads();
    }
    if (k >= T)
        return;
    const uint64_t o = min(status[2 + blockIdx.x] + s[t] - l, need);
    const uint32_t cut = (uint32_t)min<uint64_t>(l, need - o);
    if (o < need && need - 1 < o + cut) // this piece holds start need - 1
        status[1] = a.lists[(size_t)k * a.cap + beg[k] + (need - 1 - o)];
    off[k] = o;
    len[k] = cut;
}

void launch_fe_stitch(const FeWalkArgs &a, uint64_t need, uint32_t *begin, uint32_t *len, uint64_t *off,
                      uint64_t *status, hipStream_t st)
{
    const uint32_t nblocks = (a.nwalkers + FE_STITCH_BLOCK - 1) / FE_STITCH_BLOCK; // <= FE_STITCH_THREADS
    hipLaunchKernelGGL(fe_stitch_pieces_kernel, dim3(nblocks), dim3(FE_STITCH_BLOCK), 0, st, a, begin, len, status);
    hipLaunchKernelGGL(fe_stitch_scan_kernel, dim3(1), dim3(FE_STITCH_THREADS), 0, st, nblocks, need, status);
    hipLaunchKernelGGL(fe_stitch_cut_kernel, dim3(nblocks), dim3(FE_STITCH_BLOCK), 0, st, a, need, begin, len, off,
                       status);
}

// One block per piece.
__global__ __launch_bounds__(256) void fe_scatter_kernel(FeWalkArgs a, const uint32_t *begin, const uint32_t *len,
                                                         const uint64_t *off, uint64_t *out)
{
    const uint32_t k = blockIdx.x;
    const uint64_t *src = a.lists + (size_t)k * a.cap + begin[k];
    uint64_t *dst = out + off[k];
    for (uint32_t i = threadIdx.x; i < len[k]; i += blockDim.x)
        dst[i] = src[i];
}

void launch_fe_scatter(const FeWalkArgs &a, const uint32_t *begin, const uint32_t *len, const uint64_t *off,
                       uint64_t *out, hipStream_t st)
{
    hipLaunchKernelGGL(fe_scatter_kernel, dim3(a.nwalkers), dim3(256), 0, st, a, begin, len, off, out);
}

// Events of batch b_first + k.
__device__ __forceinline__ uint32_t fe_batch_events(const FeBatchWalkArgs &a, uint64_t k)
{
    const uint64_t left = a.num_events - (a.b_first + k) * a.batch_size;
    return left < a.batch_size ? (uint32_t)left : a.batch_size;
}

// The batch's stored stream as a walk's view of it.
__device__ __forceinline__ FeWalkArgs fe_batch_view(const FeBatchWalkArgs &a, uint64_t k)
{
    FeWalkArgs w;
    w.buf = a.buf;
    w.kind = a.kind;
    w.acc1 = a.acc1;
    w.acc3 = a.acc3;
    w.bbase = 0;
    w.qbase = 0;
    w.qlimit = (k + 1) * a.piece_stride;
    return w;
}

// Walker s of batch k (thread k * S + s).  Every step records a start and
// then consumes at least 15 positions, and the loop stops at the segment end
// plus the overlap, the piece's end or the list capacity, so it always ends.
// The tests come from the accept bits through LDS windows (every kind).
__global__ __launch_bounds__(BW_THREADS) void fe_seg_bit_walk_kernel(FeBatchWalkArgs a)
{
    __shared__ uint32_t win[2 * BW_THREADS * BW_ROW];
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nb * a.S)
        return;
    const uint64_t k = t / a.S;
    const uint32_t s = (uint32_t)(t % a.S);
    const uint64_t base = k * a.piece_stride, end = base + a.piece_stride;
    const uint64_t q = s == 0 ? base + (a.p0 ? (uint64_t)a.p0[k] : 0) : base + (uint64_t)s * a.seg;
    const uint64_t stop = min(base + (uint64_t)(s + 1) * a.seg + a.overlap, end);
    BitWin v;
    v.w1 = win + threadIdx.x * BW_ROW;
    v.w3 = win + (BW_THREADS + threadIdx.x) * BW_ROW;
    uint64_t *list = a.lists + t * a.cap;
    a.counts[t] = a.kind >= 2 ? bit_walk_run(XoBitWalk{a.acc1, a.acc3, 0, end}, v, q, stop, list, a.cap)
                              : bit_walk_run(RanfBitWalk{a.acc1, a.acc3, 0, end, 0}, v, q, stop, list, a.cap);
}

// Walker s's first start that walker s - 1 of the same batch also reached
// past segment s's start, as fe_join_kernel.
__global__ __launch_bounds__(256) void fe_seg_join_kernel(FeBatchWalkArgs a)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nb * a.S)
        return;
    const uint64_t k = t / a.S;
    const uint32_t s = (uint32_t)(t % a.S);
    if (s == 0) {
        a.join_prev[t] = a.join_next[t] = 0;
        return;
    }
    const uint64_t *pa = a.lists + (t - 1) * a.cap, *pb = a.lists + t * a.cap;
    const uint32_t na = a.counts[t - 1], nb = a.counts[t];
    const uint64_t qs = k * a.piece_stride + (uint64_t)s * a.seg;
    uint32_t lo = 0, hi = na; // first index of list s - 1 with a start >= qs
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (pa[mid] < qs)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t i = lo, j = 0;
    int32_t ji = -1, jj = -1;
    while (i < na && j < nb) {
        if (pa[i] == pb[j]) {
            ji = (int32_t)i;
            jj = (int32_t)j;
            break;
        }
        if (pa[i] < pb[j])
            ++i;
        else
            ++j;
    }
    a.join_prev[t] = ji;
    a.join_next[t] = jj;
}

// One wavefront per batch: the true walk is walker 0's list from its first
// start, then each walker's from where its predecessor met it.  Its first n
// starts go to the batch's slots of evq, each piece copied by the 64 lanes
// together; seq[k] = 1 (nothing written) when the true walk leaves a list
// before meeting the next one, or the lists end, before n starts.  The piece
// bounds are the same in every lane (uniform loads), so the loops are too.
__global__ __launch_bounds__(64) void fe_seg_assemble_kernel(FeBatchWalkArgs a)
{
    const uint64_t k = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (k >= a.nb)
        return;
    const uint32_t n = fe_batch_events(a, k);
    // check first, so that a batch left to the one-thread walk gets no partial starts
    uint32_t begin = 0;
    uint64_t have = 0;
    bool ok = false;
    for (uint32_t s = 0; s < a.S; ++s) {
        const uint64_t w = k * a.S + s;
        const int32_t jp = s + 1 < a.S ? a.join_prev[w + 1] : -1;
        const bool meets = jp >= 0 && (uint32_t)jp >= begin;
        const uint32_t end = meets ? (uint32_t)jp : a.counts[w];
        have += end > begin ? end - begin : 0;
        if (have >= n) {
            ok = true;
            break;
        }
        if (!meets)
            break;
        begin = (uint32_t)a.join_next[w + 1];
    }
    if (lane == 0)
        a.seq[k] = ok ? 0u : 1u;
    if (!ok)
        return;
    uint64_t *out = a.evq + k * a.batch_size;
    uint64_t off = 0;
    begin = 0;
    for (uint32_t s = 0; s < a.S && off < n; ++s) {
        const uint64_t w = k * a.S + s;
        const int32_t jp = s + 1 < a.S ? a.join_prev[w + 1] : -1;
        const bool meets = jp >= 0 && (uint32_t)jp >= begin;
        const uint32_t end = meets ? (uint32_t)jp : a.counts[w];
        const uint64_t *list = a.lists + w * a.cap;
        const uint32_t cnt = end > begin ? (uint32_t)min<uint64_t>(end - begin, n - off) : 0u;
        for (uint32_t i = lane; i < cnt; i += 64)
            out[off + i] = list[begin + i];
        off += cnt;
        if (meets)
            begin = (uint32_t)a.join_next[w + 1];
    }
}

// One thread per batch that the joined walk did not cover: its walk from its
// piece's first position.  Each event consumes at least 15 positions and the
// walk stops at the piece's end, so the loop ends.
__global__ __launch_bounds__(64) void fe_seq_walk_kernel(FeBatchWalkArgs a)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nb || !a.seq[k])
        return;
    const uint32_t n = fe_batch_events(a, k);
    const uint64_t base = k * a.piece_stride;
    const FeWalkArgs w = fe_batch_view(a, k);
    uint64_t q = base + (a.p0 ? (uint64_t)a.p0[k] : 0);
    const uint64_t q0 = q;
    uint64_t *out = a.evq + k * a.batch_size;
    uint32_t e = 0;
    for (; e < n; ++e) {
        out[e] = q;
        if (!fe_walk_any(w, q))
            break;
    }
    if (e < n) {
        *a.overflow = 1u;
        for (; e < n; ++e)
            out[e] = q0;
    }
}

void launch_fe_walk_batches(const FeBatchWalkArgs &a, hipStream_t st)
{
    const uint64_t walkers = a.nb * a.S;
    const unsigned wb = (unsigned)((walkers + 255) / 256), bb = (unsigned)((a.nb + 63) / 64);
    hipLaunchKernelGGL(fe_seg_bit_walk_kernel, dim3((unsigned)((walkers + BW_THREADS - 1) / BW_THREADS)),
                       dim3(BW_THREADS), 0, st, a);
    hipLaunchKernelGGL(fe_seg_join_kernel, dim3(wb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(fe_seg_assemble_kernel, dim3((unsigned)a.nb), dim3(64), 0, st, a);
    hipLaunchKernelGGL(fe_seq_walk_kernel, dim3(bb), dim3(64), 0, st, a);
}

// ---- faster-evgen on the xoshiro256+ stream, jump seeding --------------------
// Each piece's first state J^b s0, one thread per piece (divergent bit tests;
// the matrix columns still come in with uniform scalar loads).
__global__ __launch_bounds__(64) void xo_piece_state_kernel(XoPieceArgs a)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nb)
        return;
    const uint64_t b = a.b_first + k;
    uint64_t s[4] = {a.s0[0], a.s0[1], a.s0[2], a.s0[3]};
    for (int j = 0; j < a.kb; ++j)
        if ((b >> j) & 1ull)
            gf2_apply(a.bt + 1024 * j, s);
#pragma unroll
    for (int w = 0; w < 4; ++w)
        a.states[4 * k + w] = s[w];
}

struct Xo256Gen {
    uint64_t s[4];
    __device__ __forceinline__ double next() { return xo_uniform(s); }
};
struct Xo128Gen {
    uint32_t s[4];
    __device__ __forceinline__ float next() { return xo128_uniform(s); }
};

// One wavefront per piece.  Lane l expands positions [l * pt, (l + 1) * pt)
// of the piece from its state skipped l * pt draws, one 128-byte line of
// values (16 f64 or 32 f32) at a time, with the draws of positions
// p + 1 .. p + 3 in hand when position p is stored, so that p's accept bits
// (with p + 1 and with p + 3) are known; 3 draws past the lane's range
// complete its last bits.  Each block goes out through LDS: a store then
// writes whole lines (4 lanes' blocks of f64, 2 of f32) instead of 64
// scattered values.  The bit words go out 4 at a time (128 positions) per
// lane.
template <class T, class Gen>
__device__ __forceinline__ void xo_expand_piece(Gen g, T *piece, uint32_t *acc1, uint32_t *acc3, uint32_t pt)
{
    constexpr int B = 128 / sizeof(T), ROW = B + 1; // values per block (a 128-byte line), padded LDS rows
    constexpr int NB = 128 / B;        // blocks per group of 4 bit words
    __shared__ T st[64 * ROW];
    __shared__ uint4 sw[2 * 64];       // the lane's bit words of the current 128 positions
    const uint32_t lane = threadIdx.x;
    T *row = st + lane * ROW;
    uint32_t *m1 = reinterpret_cast<uint32_t *>(sw + lane), *m3 = reinterpret_cast<uint32_t *>(sw + 64 + lane);
    uint4 *w1 = reinterpret_cast<uint4 *>(acc1 + (uint64_t)lane * (pt / 32));
    uint4 *w3 = reinterpret_cast<uint4 *>(acc3 + (uint64_t)lane * (pt / 32));
    T u0 = g.next(), u1 = g.next(), u2 = g.next(); // positions p, p + 1, p + 2
    for (uint32_t grp = 0; grp < pt / 128; ++grp) {
#pragma unroll 1
        for (int b = 0; b < NB; ++b) {
            uint32_t b1 = 0, b3 = 0;
#pragma unroll
            for (int i = 0; i < B; ++i) {
                const T u3 = g.next();
                const T x0 = (T)2. * u0 - (T)1.;
                b1 |= disc_accept(x0, (T)2. * u1 - (T)1.) << i;
                b3 |= disc_accept(x0, (T)2. * u3 - (T)1.) << i;
                row[i] = u0;
                u0 = u1;
                u1 = u2;
                u2 = u3;
            }
            const int wi = b * B / 32, sh = b * B % 32;
            m1[wi] = sh ? m1[wi] | (b1 << sh) : b1;
            m3[wi] = sh ? m3[wi] | (b3 << sh) : b3;
            __syncthreads();
            // each store: 64 / B lanes' blocks, B consecutive positions each
            const uint64_t pos = (uint64_t)grp * 128 + (uint64_t)b * B;
#pragma unroll 4
            for (int r = 0; r < B; ++r) { // streaming stores of whole lines: a later kernel reads them
                const uint32_t src = r * (64 / B) + lane / B, i = lane % B;
                __builtin_nontemporal_store(st[src * ROW + i], &piece[(uint64_t)src * pt + pos + i]);
            }
            __syncthreads();
        }
        w1[grp] = sw[lane];
        w3[grp] = sw[64 + lane];
    }
}

// Skip of the lane's state to its first draw: M^(lane * pt) as the product
// of M^(pt 2^j) over the set bits j of the lane index, six matrices that every
// lane reads alike.
__global__ __launch_bounds__(64) void xo_expand_kernel(XoPieceArgs a)
{
    const uint64_t k = blockIdx.x;
    Xo256Gen g;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        g.s[w] = a.states[4 * k + w];
#pragma unroll 1
    for (int j = 0; j < 6; ++j)
        if ((threadIdx.x >> j) & 1u)
            gf2_apply(a.lt + 1024 * j, g.s);
    const uint64_t words = a.piece_len / 32;
    xo_expand_piece<double>(g, a.xbuf + k * a.piece_len, a.acc1 + k * words, a.acc3 + k * words, a.per_thread);
}
// xoshiro128+ pieces, as the xoshiro256+ ones above.
__global__ __launch_bounds__(64) void xo128_piece_state_kernel(Xo128PieceArgs a)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nb)
        return;
    const uint64_t b = a.b_first + k;
    uint32_t s[4] = {a.s0[0], a.s0[1], a.s0[2], a.s0[3]};
    for (int j = 0; j < a.kb; ++j)
        if ((b >> j) & 1ull)
            gf2_apply128(a.bt + 512 * j, s);
#pragma unroll
    for (int w = 0; w < 4; ++w)
        a.states[4 * k + w] = s[w];
}

__global__ __launch_bounds__(64) void xo128_expand_kernel(Xo128PieceArgs a)
{
    const uint64_t k = blockIdx.x;
    Xo128Gen g;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        g.s[w] = a.states[4 * k + w];
#pragma unroll 1
    for (int j = 0; j < 6; ++j)
        if ((threadIdx.x >> j) & 1u)
            gf2_apply128(a.lt + 512 * j, g.s);
    const uint64_t words = a.piece_len / 32;
    xo_expand_piece<float>(g, a.xbuf + k * a.piece_len, a.acc1 + k * words, a.acc3 + k * words, a.per_thread);
}

void launch_xo128_expand(const Xo128PieceArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(xo128_piece_state_kernel, dim3((unsigned)((a.nb + 63) / 64)), dim3(64), 0, st, a);
    hipLaunchKernelGGL(xo128_expand_kernel, dim3((unsigned)a.nb), dim3(64), 0, st, a);
}

void launch_xo_expand(const XoPieceArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(xo_piece_state_kernel, dim3((unsigned)((a.nb + 63) / 64)), dim3(64), 0, st, a);
    hipLaunchKernelGGL(xo_expand_kernel, dim3((unsigned)a.nb), dim3(64), 0, st, a);
}


// seeded_new (ranf.rs:36-66) for jump mode, one thread per batch.
__global__ void ranf_seed_kernel(int32_t seed0, uint64_t b_first, uint64_t nb, int32_t *origin, int32_t *p0)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb)
        return;
    const uint64_t b = b_first + i;
    int32_t x[56];
    const int32_t seed = (int32_t)((uint32_t)seed0 + (uint32_t)(123456ull * b));
#pragma unroll
    for (int k = 0; k < 56; ++k)
        x[k] = 0;
    x[55] = seed;
    int32_t jj = seed, kk = 1;
    for (int t = 1; t < 55; ++t) {
        const int ii = (21 * t) % 55;
        // dynamic index into a private array: stays in scratch-free registers
        // only if unrolled; the loop is short and runs once per batch.
        x[ii] = kk;
        kk = (int32_t)((uint32_t)jj - (uint32_t)kk);
        if (kk < 0)
            kk = (int32_t)((uint32_t)kk + (uint32_t)MODULO);
        jj = x[ii];
    }
    for (int rr = 0; rr < 10; ++rr) {
        for (int t = 1; t < 25; ++t) {
            x[t] -= x[t + 31];
            if (x[t] < 0)
                x[t] += MODULO;
        }
        for (int t = 25; t < 56; ++t) {
            x[t] -= x[t - 24];
            if (x[t] < 0)
                x[t] += MODULO;
        }
    }
    for (int k = 0; k < 56; ++k)
        origin[56 * i + k] = x[k];
    p0[i] = 0;
}

void launch_ranf_seed(int32_t seed0, uint64_t b_first, uint64_t nb, int32_t *origin, int32_t *p0, hipStream_t st)
{
    const unsigned blocks = (unsigned)((nb + 63) / 64);
    hipLaunchKernelGGL(ranf_seed_kernel, dim3(blocks), dim3(64), 0, st, seed0, b_first, nb, origin, p0);
}

// ---- in-order merge --------------------------------------------------------
// The fold must stay a left fold in batch order (its rounding is the
// reference's), so it is one dependent chain of adds per field.  Tiles of
// accumulators are staged through LDS with coalesced loads by the whole
// workgroup; 12 lanes then run the 12 chains out of LDS, so the chain pays
// the add latency, not the HBM latency.  The event count is an integer sum
// and is reduced in any order.
constexpr int MERGE_TILE = 256;
constexpr int MERGE_THREADS = 256;

// In-order left fold of the per-batch accumulators (resacc.rs:133-139): the
// 12 sums are 12 chains of dependent adds that must run in batch order, so
// their length (nb adds) is the floor.  Wave 0 (lanes 0..11) runs the chains
// over one LDS tile while waves 1..3 load the next tile from HBM into the
// other buffer; the selected counts are integers and are summed by the loaders
// in any order.
template <class T>
__global__ __launch_bounds__(MERGE_THREADS) void merge_kernel(const DevAcc *in, uint64_t nb, DevAcc *out)
{
    __shared__ double tile[2][12][MERGE_TILE + 1]; // +1: the 12 chains hit distinct banks
    __shared__ unsigned long long cnt;
    const int tid = threadIdx.x;
    if (tid == 0)
        cnt = 0;
    const uint64_t *words = reinterpret_cast<const uint64_t *>(in); // 13 words per accumulator
    const uint64_t ntiles = (nb + MERGE_TILE - 1) / MERGE_TILE;
    unsigned long long my_cnt = 0;
    // tile k -> buffer k & 1, loaded by threads [first, MERGE_THREADS): each
    // thread issues all its loads before its first LDS store, so a tile costs
    // one memory latency rather than one per word.
    constexpr int PER = (13 * MERGE_TILE + (MERGE_THREADS - 64) - 1) / (MERGE_THREADS - 64);
    auto load = [&](uint64_t k, int first) {
        const uint64_t base = k * MERGE_TILE;
        const uint32_t m = (uint32_t)((nb - base) < (uint64_t)MERGE_TILE ? (nb - base) : (uint64_t)MERGE_TILE);
        const uint32_t nw = 13 * m, step = (uint32_t)(MERGE_THREADS - first);
        const uint64_t *src = words + 13 * base;
        double(*dst)[MERGE_TILE + 1] = tile[k & 1];
        uint64_t v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const uint32_t f = (uint32_t)(tid - first) + (uint32_t)i * step;
            v[i] = f < nw ? src[f] : 0ull;
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const uint32_t f = (uint32_t)(tid - first) + (uint32_t)i * step;
            if (f < nw) {
                const uint32_t e = f / 13, w = f - 13 * e;
                if (w == 0)
                    my_cnt += v[i];
                else
                    dst[w - 1][e] = __longlong_as_double((long long)v[i]);
            }
        }
    };
    if (ntiles > 0)
        load(0, 0);
    __syncthreads();
    T t = (T)0;
    for (uint64_t k = 0; k < ntiles; ++k) {
        if (tid >= 64) {
            if (k + 1 < ntiles)
                load(k + 1, 64);
        } else if (tid < 12) {
            const uint64_t base = k * MERGE_TILE;
            const uint32_t m = (uint32_t)((nb - base) < (uint64_t)MERGE_TILE ? (nb - base) : (uint64_t)MERGE_TILE);
            const double *row = tile[k & 1][tid];
            uint32_t e = 0;
            if (k == 0) {
                t = (T)row[0];
                e = 1;
            }
            // Software-pipelined: the LDS reads of the next 16 values are in
            // flight while the current 16 go through the add chain.
            if (e + 16 <= m) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    v[u] = row[e + u];
                for (e += 16; e + 16 <= m; e += 16) {
                    double w[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        w[u] = row[e + u];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        t = t + (T)v[u];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        v[u] = w[u];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    t = t + (T)v[u];
            }
            for (; e < m; ++e)
                t = t + (T)row[e];
        }
        __syncthreads();
    }
    atomicAdd(&cnt, my_cnt);
    __syncthreads();
    if (tid < 12)
        reinterpret_cast<double *>(out)[1 + tid] = (double)t;
    if (tid == 0)
        out->selected = cnt;
}

void launch_merge(const void *per_batch, uint64_t nb, void *merged, hipStream_t st, int f32)
{
    if (f32)
        hipLaunchKernelGGL(merge_kernel<float>, dim3(1), dim3(MERGE_THREADS), 0, st,
                           reinterpret_cast<const DevAcc *>(per_batch), nb, reinterpret_cast<DevAcc *>(merged));
    else
        hipLaunchKernelGGL(merge_kernel<double>, dim3(1), dim3(MERGE_THREADS), 0, st,
                           reinterpret_cast<const DevAcc *>(per_batch), nb, reinterpret_cast<DevAcc *>(merged));
}

// ---- all-f32 mode --------------------------------------------------------------
// One wavefront per batch (grid-stride).  Each lane takes a contiguous run of
// the batch's events (one skip-ahead of the xoshiro128+ stream per run; with
// faster-evgen every 64th event instead) and writes each event's m[5] and
// cut decision; the batch sums are then formed
// in event order, as the reference's sequential loop does in f32
// (resacc.rs:121-129), where a tree sum would round differently.
// Events that pass the cuts are queued in LDS (momenta and event index) and
// their helicity sums computed 64 at a time on dense lanes, as in the f64
// event loop; each event's m[5] and cut flag go to its own slot of a.ev.
template <int SOURCE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) void f32_events_kernel(F32Args a)
{
    __shared__ float queue[12][QCAP];
    __shared__ uint32_t qidx[QCAP];
    const uint32_t lane = threadIdx.x;
    const uint64_t lanemask_lt = (1ull << lane) - 1ull;
    // helicity sums of the queued event in `slot`, stored to its output slot
    auto flush_one = [&](uint32_t slot) {
        float pq[3][4], m[5];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                pq[i][j] = queue[4 * i + j][slot];
        const uint32_t idx = qidx[slot];
        matrix_elements_f32(a.P, pq, m);
        float *o = a.ev + 6 * (size_t)idx;
#pragma unroll
        for (int c = 0; c < 5; ++c)
            o[c] = m[c];
        o[5] = 1.f;
    };
    uint32_t st[4] = {0, 0, 0, 0}; // xoshiro128+ state at this lane's chunk of the previous batch
    bool have = false;
    for (uint64_t lb = blockIdx.x; lb < a.nb; lb += gridDim.x) {
        const uint64_t b = a.b_first + lb;
        const uint64_t left = a.num_events - b * a.batch_size;
        const uint32_t n = left < a.batch_size ? (uint32_t)left : a.batch_size;
        const uint32_t q = n / 64, r = n % 64;
        const uint32_t len = q + (lane < r ? 1u : 0u);
        // a contiguous run of events per lane (one skip-ahead of the stream
        // per run) for the sources that draw 12 values per event; with
        // faster-evgen the events lane, lane + 64, ... (their draws lie side
        // by side across the wavefront)
        constexpr bool FE = SOURCE == 2 || SOURCE == 3;
        const uint32_t e0 = FE ? lane : lane * q + (lane < r ? lane : r);
        const uint32_t estep = FE ? 64u : 1u;
        uint32_t s[4] = {a.s0[0], a.s0[1], a.s0[2], a.s0[3]};
        if (SOURCE == 0) {
            // M^(12 e0) B^b s0; the next batch of the grid-stride loop is B^grid
            // times the same (all factors are powers of the step matrix, so
            // they commute), unless its event split differs (the last batch)
            if (have && n == a.batch_size && (int)gridDim.x == a.stride_grid) {
                gf2_apply128(a.stride, st);
            } else {
                st[0] = a.s0[0];
                st[1] = a.s0[1];
                st[2] = a.s0[2];
                st[3] = a.s0[3];
                for (int k = 0; k < a.kb; ++k)
                    if ((b >> k) & 1ull) // uniform across the wavefront
                        gf2_apply128(a.bt + 512 * k, st);
                const uint64_t off = 12ull * e0;
                for (int k = 0; k < a.km; ++k)
                    if ((off >> k) & 1ull)
                        gf2_apply128(a.mt + 512 * k, st);
            }
            have = n == a.batch_size;
            s[0] = st[0];
            s[1] = st[1];
            s[2] = st[2];
            s[3] = st[3];
        }
        const uint32_t iters = q + (r ? 1u : 0u);
        uint32_t qn = 0; // queued events (uniform)
        // faster-evgen: the lane's next event start, loaded an event ahead
        uint64_t qnext = FE && len > 0 ? a.evq[lb * a.batch_size + e0] : 0;
        for (uint32_t i = 0; i < iters; ++i) {
            const uint64_t e = lb * a.batch_size + e0 + (uint64_t)i * estep;
            bool pass = false;
            float p[3][4];
            const uint64_t qk = qnext;
            if (FE && i + 1 < len)
                qnext = a.evq[e + estep];
            if (i < len && FE) {
                const RerollBits rb{a.acc1, a.bbase, a.bstride ? (lb + 1) * a.bstride : a.blimit};
                if constexpr (SOURCE == 2) {
                    RanfWalkF32 w{a.draws, a.qbase, qk, rb};
                    pass = generate_faster_f32(a.P, w, p);
                } else {
                    Xo128WalkF32 w{a.x32draws, a.qbase, qk, rb};
                    pass = generate_faster_f32(a.P, w, p);
                }
                if (!pass) {
                    float *o = a.ev + 6 * e;
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        o[c] = 0.f;
                }
            } else if (i < len) {
                float u[12];
                if (SOURCE == 0) {
#pragma unroll
                    for (int j = 0; j < 12; ++j)
                        u[j] = xo128_uniform(s);
                } else {
                    const int32_t *d = a.draws + 12 * e;
#pragma unroll
                    for (int j = 0; j < 12; ++j)
                        u[j] = (float)d[j] * 1e-9f; // (n as f32) * INV_MODULO (ranf.rs:99)
                }
                pass = generate_f32(a.P, u, p);
                if (!pass) {
                    float *o = a.ev + 6 * e;
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        o[c] = 0.f;
                }
            }
            const uint64_t mask = __ballot(pass);
            if (pass) {
                const uint32_t slot = qn + (uint32_t)__popcll(mask & lanemask_lt);
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        queue[4 * k + j][slot] = p[k][j];
                qidx[slot] = (uint32_t)e;
            }
            qn += (uint32_t)__popcll(mask);
            if (qn >= 64) {
                __syncthreads();
                flush_one(qn - 64 + lane);
                __syncthreads();
                qn -= 64;
            }
        }
        if (qn > 0) { // partial flush
            __syncthreads();
            if (lane < qn)
                flush_one(lane);
        }
        __syncthreads();
    }
}

constexpr int F32_TILE = 256;

// In-order f32 batch sums (resacc.rs:121-129 with Float = f32).  For each
// tile of events all 64 lanes form the 12 summands (m_c, m_c^2, w, w^2) in
// parallel, writing -0.0 for a rejected event (x + (-0.0) == x for every x,
// so a rejected event changes no sum, as when the reference skips it); then
// lanes 0..11 add their summand column in event order, which is the order
// that fixes the f32 rounding.  Selected events are counted with ballots.
__global__ __launch_bounds__(64) void f32_accumulate_kernel(F32Args a)
{
    __shared__ float terms[12][F32_TILE + 1];
    const uint32_t lane = threadIdx.x;
    DevAcc *out = reinterpret_cast<DevAcc *>(a.out);
    for (uint64_t lb = blockIdx.x; lb < a.nb; lb += gridDim.x) {
        const uint64_t b = a.b_first + lb;
        const uint64_t left = a.num_events - b * a.batch_size;
        const uint32_t n = left < a.batch_size ? (uint32_t)left : a.batch_size;
        const float *ev = a.ev + 6 * lb * a.batch_size;
        float t = 0.f; // lane j < 12: sum j of (spm2[5], vars[5], sigma, variance)
        uint32_t cnt = 0;
        for (uint32_t base = 0; base < n; base += F32_TILE) {
            const uint32_t mm = n - base < (uint32_t)F32_TILE ? n - base : (uint32_t)F32_TILE;
            __syncthreads(); // the previous tile has been summed
            for (uint32_t e = lane; e < mm; e += 64) {
                const float *x = ev + 6 * (size_t)(base + e);
                float m[5];
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    m[c] = x[c];
                const bool keep = x[5] != 0.f;
                cnt += (uint32_t)__popcll(__ballot(keep));
                const float w = weight_f32(a.P, m);
                const float z = -0.f;
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    terms[c][e] = keep ? m[c] : z;
                    terms[5 + c][e] = keep ? m[c] * m[c] : z;
                }
                terms[10][e] = keep ? w : z;
                terms[11][e] = keep ? w * w : z;
            }
            __syncthreads();
            if (lane < 12) {
                const float *row = terms[lane];
                uint32_t e = 0;
                for (; e + 16 <= mm; e += 16) {
                    float v[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        v[u] = row[e + u];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        t = t + v[u];
                }
                for (; e < mm; ++e)
                    t = t + row[e];
            }
        }
        if (lane < 12)
            reinterpret_cast<double *>(out + lb)[1 + lane] = (double)t;
        if (lane == 0) // lane 0 runs every pass of the tile loop, so its count is complete
            out[lb].selected = cnt;
    }
}

void launch_f32_events(const F32Args &a, int grid, hipStream_t st)
{
    if (a.source == 0)
        hipLaunchKernelGGL(f32_events_kernel<0>, dim3(grid), dim3(64), 0, st, a);
    else if (a.source == 1)
        hipLaunchKernelGGL(f32_events_kernel<1>, dim3(grid), dim3(64), 0, st, a);
    else if (a.source == 2)
        hipLaunchKernelGGL(f32_events_kernel<2>, dim3(grid), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL(f32_events_kernel<3>, dim3(grid), dim3(64), 0, st, a);
}

void launch_f32_accumulate(const F32Args &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(f32_accumulate_kernel, dim3(grid), dim3(64), 0, st, a);
}

} // namespace tpk

// ---- adjoint ---------------------------------------------------------------
#include "tp_adjoint.h"
#include "tp_adjstore.h"

namespace tpk {

// One 64-lane workgroup per 64 consecutive points.  The sweeps run
// convergently (a cut point's lane runs them too and hands zeros on), so the
// sink can synchronise the workgroup: it stages F and dF/du in LDS for the
// whole pass and dF/dp one sweep at a time, and the stores go out with
// consecutive lanes on consecutive addresses (tp_adjstore.h), whole cache
// lines at a time for F and dF/du.
// Stage elements e = lane + 64 k, k < K, of a full workgroup (np = 64) to
// put(e, value): a fixed trip count, so the LDS loads issue back to back
// instead of one load-wait-store round trip per element.
template <int K, class T, class Put> __device__ __forceinline__ void flush_full(const T *stage, uint32_t lane, Put &&put)
{
#pragma unroll
    for (int k = 0; k < K; ++k)
        put(lane + tpa::LANES * k, stage[lane + tpa::LANES * k]);
}

// One point per lane, each sweep's values stored straight from registers:
// a store instruction writes 64 rows 480 B apart (f64), so lines reach HBM
// partly written (WRITE_SIZE 1.28x the output bytes), but nothing waits on
// them.  The f64 kernel, at one wavefront per SIMD, is fastest this way.
template <class T> __device__ __forceinline__ void adjoint_direct(const AdjArgs &a)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    const AdjParams<T> P = adj_params<T>(a.P, a.w);
    T *F = reinterpret_cast<T *>(a.F), *Gu = reinterpret_cast<T *>(a.dF_du), *Gp = reinterpret_cast<T *>(a.dF_dp);
    T u[12];
#pragma unroll
    for (int j = 0; j < 12; ++j)
        u[j] = (T)a.u_in[12 * i + j];
    // one forward pass, then one reverse sweep per coupling component
    T m[5], du[12], dp[12];
    event_adjoint_sweeps<T, 0, 5>(P, u, nullptr, m, du, dp, (T)0, (T)INFINITY,
                                  [&](int c, const T *mc, const T *duc, const T *dpc) {
                                      if (F)
                                          F[5 * i + c] = P.w * (c == 0 ? mc[0] : c == 1 ? mc[1] : c == 2 ? mc[2] : c == 3 ? mc[3] : mc[4]);
                                      if (Gu)
#pragma unroll
                                          for (int j = 0; j < 12; ++j)
                                              Gu[60 * i + 12 * c + j] = duc[j];
                                      if (Gp)
#pragma unroll
                                          for (int j = 0; j < 12; ++j)
                                              Gp[60 * i + 12 * c + j] = dpc[j];
                                  });
}

// One point per lane, rows staged in LDS and stored with consecutive lanes
// on consecutive addresses (the index arithmetic is in tp_adjstore.h):
// WRITE_SIZE 1.35x the output bytes in f32 instead of 1.72x, and 13 % less
// time.  Every lane of the workgroup reaches every barrier: lanes past n
// compute a copy of point base and store nothing, and a cut point runs the
// sweeps like the others and hands zeros to the sink.
template <class T> __device__ __forceinline__ void adjoint_staged(const AdjArgs &a, T *sF, T *sU, T *sP)
{
    const uint32_t lane = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * tpa::LANES;
    if (base >= a.n) // whole workgroup
        return;
    const uint32_t np = tpa::points(a.n, base);
    const bool full = np == tpa::LANES;
    const uint64_t i = base + lane;
    const bool valid = lane < np;
    const AdjParams<T> P = adj_params<T>(a.P, a.w);
    T *F = reinterpret_cast<T *>(a.F), *Gu = reinterpret_cast<T *>(a.dF_du), *Gp = reinterpret_cast<T *>(a.dF_dp);
    T u[12];
#pragma unroll
    for (int j = 0; j < 12; ++j)
        u[j] = (T)a.u_in[12 * (valid ? i : base) + j]; // lanes past n: a copy of point base, never stored
    // one forward pass, then one reverse sweep per coupling component
    T m[5], du[12], dp[12];
    event_adjoint_sweeps<T, 0, 5, true>(
        P, u, nullptr, m, du, dp, (T)0, (T)INFINITY, [&](int c, const T *mc, const T *duc, const T *dpc) {
            if (F)
                sF[tpa::slot_f(lane, c)] = P.w * (c == 0 ? mc[0] : c == 1 ? mc[1] : c == 2 ? mc[2] : c == 3 ? mc[3] : mc[4]);
            if (Gu)
#pragma unroll
                for (int j = 0; j < 12; ++j)
                    sU[tpa::slot_du(lane, c, j)] = duc[j];
            if (Gp) {
                __syncthreads(); // the previous sweep's dF/dp rows are out
#pragma unroll
                for (int j = 0; j < 12; ++j)
                    sP[tpa::slot_dp(lane, j)] = dpc[j];
                __syncthreads();
                for (uint32_t e = lane; e < 12 * np; e += tpa::LANES)
                    Gp[tpa::dst_dp(base, e, c)] = sP[e];
            }
        });
    __syncthreads();
    if (Gu) {
        if (full)
            flush_full<60>(sU, lane, [&](uint32_t e, T v) { Gu[tpa::dst_du(base, e)] = v; });
        else
            for (uint32_t e = lane; e < 60 * np; e += tpa::LANES)
                Gu[tpa::dst_du(base, e)] = sU[e];
    }
    if (F) {
        if (full)
            flush_full<5>(sF, lane, [&](uint32_t e, T v) { F[tpa::dst_f(base, e)] = v; });
        else
            for (uint32_t e = lane; e < 5 * np; e += tpa::LANES)
                F[tpa::dst_f(base, e)] = sF[e];
    }
}

// The stream's points, ADJ_GEN_POINTS consecutive ones per thread from one
// skip-ahead (12 draws per point, rand's Standard f64).
__global__ __launch_bounds__(64) void adjoint_points_kernel(AdjArgs a)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t first = t * ADJ_GEN_POINTS;
    if (first >= a.n)
        return;
    uint64_t s[4] = {a.s0[0], a.s0[1], a.s0[2], a.s0[3]};
    const uint64_t off = 12ull * first;
    for (int k = 0; k < a.km; ++k)
        if ((off >> k) & 1ull)
            gf2_apply(a.mt + 1024 * k, s);
    const uint64_t last = first + ADJ_GEN_POINTS < a.n ? first + ADJ_GEN_POINTS : a.n;
    double *out = a.u_gen + 12 * first;
    for (uint64_t k = 0; k < 12 * (last - first); ++k)
        out[k] = xo_uniform(s);
}

// f64: 256 VGPRs + 218 AGPRs, one wavefront per SIMD (two would spill
// hundreds of bytes per lane); with the staged store (39.4 KB of LDS) it ran
// 2.5 % slower than with direct stores, so it stores directly.
// f32: 255 VGPRs, two wavefronts per SIMD (waves_per_eu(2, 2)), no spill;
// 19.7 KB of stage, eight workgroups per CU.
__global__ __launch_bounds__(64) void adjoint_kernel_f64(AdjArgs a) { adjoint_direct<double>(a); }
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void adjoint_kernel_f32(AdjArgs a)
{
    __shared__ float sF[5 * tpa::LANES], sU[60 * tpa::LANES], sP[12 * tpa::LANES];
    adjoint_staged<float>(a, sF, sU, sP);
}

void launch_adjoint(const AdjArgs &a0, int precision, hipStream_t st)
{
    AdjArgs a = a0;
    if (!a.u_in) {
        const uint64_t threads = (a.n + ADJ_GEN_POINTS - 1) / ADJ_GEN_POINTS;
        hipLaunchKernelGGL(adjoint_points_kernel, dim3((unsigned)((threads + 63) / 64)), dim3(64), 0, st, a);
        a.u_in = a.u_gen;
    }
    const dim3 g((unsigned)((a.n + tpa::LANES - 1) / tpa::LANES)), b(tpa::LANES);
    if (precision == 2)
        hipLaunchKernelGGL(adjoint_kernel_f32, g, b, 0, st, a);
    else
        hipLaunchKernelGGL(adjoint_kernel_f64, g, b, 0, st, a);
}

} // namespace tpk
