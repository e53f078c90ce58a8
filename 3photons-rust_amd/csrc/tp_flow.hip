// tp_flow.hip -- bf16-MFMA kernels of the RealNVP proposal (see tp_flow.h).
//
// Matrix products use v_mfma_f32_16x16x16_bf16.  Activations are kept
// transposed (features x samples): for H^T = W X^T the weight tile is the A
// operand and the 16-sample activation tile the B operand, and the lane
// layout of the f32 result (lane l holds rows 4(l/16)..4(l/16)+3 of column
// l%16) is exactly the B layout the next product needs, so a whole coupling
// layer runs lane-locally in registers.  The coupling state of one sample is
// spread over the four lanes l%16 + 16 g, lane group g holding features
// 4g..4g+3 -- the same rows the s and t outputs land in.
#include "tp_flow.h"
#include "tp_adjoint.h"
#include "tp_xoshiro.h"

#include <algorithm>

// The network code below is not bound to any reference operation order:
// let the compiler fuse multiply-adds.  (The physics included above keeps
// the build's contraction-free setting; it is switched back off before the
// integrand kernel at the end of this file.)
#pragma clang fp contract(fast)

namespace tpf {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr float LOG2E = 1.4426950408889634f, LN2 = 0.6931471805599453f;

__device__ __forceinline__ uint16_t f2bf(float x) // round to nearest even (v_cvt_pk_bf16_f32)
{
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ s4 pack4(f4 v)
{
    const bf4 b = {(__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
    return __builtin_bit_cast(s4, b);
}
__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// gfx950's 16x16x32 product (twice the K of the one above per instruction):
// lane l supplies row / column l % 16 and eight K values of group l / 16.
__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}
// The B operand of a K = 32 product from two 16-row activation tiles: lane
// group g holds rows 4g..4g+3 of each, the eight K values of its group.
__device__ __forceinline__ s8 cat8(s4 lo, s4 hi) { return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7); }
// Hardware exp2 / log2 / reciprocal (about 1 ulp): the network's activations
// are rounded to bf16 anyway.
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * LOG2E); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * LN2; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sigm(float x) { return frcp(1.f + fexp(-x)); }

// Hard-swish y = c min(max(c / 6 + 1/2, 0), 1) and its derivative (0 below
// -3, c / 3 + 1/2 up to 3, 1 above; PyTorch's hardswish): three full-rate
// operations where a sigmoid-based activation takes an exp2 and a reciprocal
// at a quarter of the rate (the forward-only kernels drop d).
__device__ __forceinline__ void hswish_d(float c, float &y, float &d)
{
    y = c * fminf(fmaxf(fmaf(c, 1.f / 6.f, 0.5f), 0.f), 1.f);
    d = c < -3.f ? 0.f : (c > 3.f ? 1.f : fmaf(c, 1.f / 3.f, 0.5f));
}
__device__ __forceinline__ float ftanh(float x) { return 1.f - 2.f * frcp(fexp(2.f * x) + 1.f); }
__device__ __forceinline__ float softplus(float x) { return fmaxf(x, 0.f) + flog(1.f + fexp(-fabsf(x))); }

// ---- LDS image of one layer (bf16, padded row strides against bank conflicts)
// W2 and W3 (K = 64) feed K = 32 products: their columns are stored in the
// order those products read them.  Within each 32-column half c, lane group g
// reads the eight stored columns 32c + 8g .. 32c + 8g + 7, which hold the
// original columns 32c + 4g + (0..3) and 32c + 16 + 4g + (0..3): the rows
// 4g..4g+3 of the two 16-row activation tiles 2c and 2c+1 that the lane
// holds (cat8).  Row strides of 72 keep each lane's 16-byte read aligned.
constexpr int S_W1 = 20, S_W2 = 72, S_W3 = 72, S_W1T = 68, S_W2T = 68, S_W3T = 36;
__host__ __device__ constexpr int k32_pos(int k) // stored position of original column k
{
    return 32 * (k / 32) + 8 * ((k % 16) / 4) + (k % 4) + 4 * ((k % 32) / 16);
}
constexpr int L_W1 = 0, L_W2 = L_W1 + H * S_W1, L_W3 = L_W2 + H * S_W2, L_FWD = L_W3 + O * S_W3;
constexpr int L_W1T = L_FWD, L_W2T = L_W1T + DP * S_W1T, L_W3T = L_W2T + H * S_W2T, L_ALL = L_W3T + H * S_W3T;

struct LdsLayer {
    uint16_t w[L_ALL];
    float b[H + H + O];
};

// Forward weights only (W1, W2, W3 and the biases, same offsets as LdsLayer):
// the forward and inverse kernels keep all L layers of them resident.
struct LdsFwd {
    uint16_t w[L_FWD];
    float b[H + H + O];
};

// One layer's weights and biases into LDS; TRANSPOSED (backward pass, an
// LdsLayer) also writes W1^T, W2^T, W3^T.
template <bool TRANSPOSED, class Lds> __device__ __forceinline__ void load_layer(Lds &s, const uint16_t *wl)
{
    static_assert(!TRANSPOSED || sizeof(s.w) == sizeof(uint16_t) * L_ALL, "the transposed copies need an LdsLayer");
    for (int i = threadIdx.x; i < H * DP; i += blockDim.x) {
        const int r = i / DP, c = i % DP;
        s.w[L_W1 + r * S_W1 + c] = wl[OFF_W1 + i];
        if constexpr (TRANSPOSED)
            s.w[L_W1T + c * S_W1T + r] = wl[OFF_W1 + i];
    }
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        const int r = i / H, c = i % H;
        s.w[L_W2 + r * S_W2 + k32_pos(c)] = wl[OFF_W2 + i];
        if constexpr (TRANSPOSED)
            s.w[L_W2T + c * S_W2T + r] = wl[OFF_W2 + i];
    }
    for (int i = threadIdx.x; i < O * H; i += blockDim.x) {
        const int r = i / H, c = i % H;
        s.w[L_W3 + r * S_W3 + k32_pos(c)] = wl[OFF_W3 + i];
        if constexpr (TRANSPOSED)
            s.w[L_W3T + c * S_W3T + r] = wl[OFF_W3 + i];
    }
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        s.b[i] = bf2f(wl[OFF_B1 + i]);
        s.b[H + i] = bf2f(wl[OFF_B2 + i]);
    }
    for (int i = threadIdx.x; i < O; i += blockDim.x)
        s.b[2 * H + i] = bf2f(wl[OFF_B3 + i]);
}

// A operand of a 16x16 tile of a row-major LDS matrix: rows 16 mb .., cols 16 kb ..
__device__ __forceinline__ s4 atile(const uint16_t *m, int stride, int mb, int kb)
{
    const int lane = threadIdx.x & 63;
    const int r = 16 * mb + (lane & 15), c = 16 * kb + 4 * (lane >> 4);
    return *reinterpret_cast<const s4 *>(m + r * stride + c);
}

// A operand of a K = 32 product: rows 16 mb .., stored columns 32 c + 8 g ..
__device__ __forceinline__ s8 atile8(const uint16_t *m, int stride, int mb, int c)
{
    const int lane = threadIdx.x & 63;
    const int r = 16 * mb + (lane & 15), col = 32 * c + 8 * (lane >> 4);
    return *reinterpret_cast<const s8 *>(m + r * stride + col);
}

// Bias values of rows r0 .. r0 + 3 (the lane's rows of an output tile).
template <class Lds> __device__ __forceinline__ f4 bias4(const Lds &s, int r0)
{
    return *reinterpret_cast<const f4 *>(s.b + r0);
}

struct MlpFwd {
    s4 a;          // bf16 input (features 4g..4g+3)
    f4 h1d[4];     // hard-swish'(pre-activation), 4 row blocks (backward only)
    s4 h1[4];
    f4 h2d[4];
    s4 h2[4];
    f4 o0, o1;     // s raw (rows 0..15), t raw (rows 16..31)
};

template <class Lds> __device__ __forceinline__ void mlp_forward(const Lds &s, f4 a_in, MlpFwd &r)
{
    const int g = (threadIdx.x & 63) >> 4;
    r.a = pack4(a_in);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const f4 c = mfma(atile(s.w + L_W1, S_W1, mb, 0), r.a, bias4(s, 16 * mb + 4 * g));
        f4 h, d;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float y, dy;
            hswish_d(c[i], y, dy);
            h[i] = y;
            d[i] = dy;
        }
        r.h1d[mb] = d;
        r.h1[mb] = pack4(h);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f4 c = bias4(s, H + 16 * mb + 4 * g);
#pragma unroll
        for (int kc = 0; kc < 2; ++kc)
            c = mfma32(atile8(s.w + L_W2, S_W2, mb, kc), cat8(r.h1[2 * kc], r.h1[2 * kc + 1]), c);
        f4 h, d;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float y, dy;
            hswish_d(c[i], y, dy);
            h[i] = y;
            d[i] = dy;
        }
        r.h2d[mb] = d;
        r.h2[mb] = pack4(h);
    }
    f4 o0 = bias4(s, 2 * H + 4 * g), o1 = bias4(s, 2 * H + 16 + 4 * g);
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        const s8 b = cat8(r.h2[2 * kc], r.h2[2 * kc + 1]);
        o0 = mfma32(atile8(s.w + L_W3, S_W3, 0, kc), b, o0);
        o1 = mfma32(atile8(s.w + L_W3, S_W3, 1, kc), b, o1);
    }
    r.o0 = o0;
    r.o1 = o1;
}

__device__ __forceinline__ bool transformed(uint32_t cond, int j) { return j < D && !((cond >> j) & 1u); }
__device__ __forceinline__ bool conditioning(uint32_t cond, int j) { return j < D && ((cond >> j) & 1u); }

__device__ __forceinline__ f4 masked(f4 z, uint32_t cond)
{
    const int g = (threadIdx.x & 63) >> 4;
    f4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        a[i] = conditioning(cond, 4 * g + i) ? z[i] : 0.f;
    return a;
}

// Sum over the four lane groups of one sample (lanes l, l^16, l^32, l^48).
__device__ __forceinline__ float sample_sum(float x)
{
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 32);
    return x;
}

// ---- counter-stream draws ----------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t stream_u64(uint64_t key, uint64_t k)
{
    return mix64(key + (k + 1) * 0x9E3779B97F4A7C15ull);
}
// (2 m + 1) 2^-24 with m the top 23 bits: strictly inside (0, 1), so logit(v)
// is finite.
__device__ __forceinline__ float stream_v(uint64_t key, uint64_t k)
{
    return (float)(uint32_t)(stream_u64(key, k) >> 41) * 0x1.0p-23f + 0x1.0p-24f;
}
// (2 m + 1) 2^-53 with m the top 52 bits: strictly inside (0, 1).
__device__ __forceinline__ double stream_u(uint64_t key, uint64_t k)
{
    return (double)(stream_u64(key, k) >> 12) * 0x1.0p-52 + 0x1.0p-53;
}
// u = sigmoid(z), evaluated in f32 (the density is that of the f32 network;
// the point moves by at most an f32 rounding), stored as f64 and kept inside
// (0, 1): at most 1 - 2^-53 where the f32 value rounds to 1.
__device__ __forceinline__ double sigmoid_u(float z)
{
    const double u = (double)sigm(z);
    return fmin(fmax(u, 0x1.0p-126), 0x1.fffffffffffffp-1);
}

// ---- forward ----------------------------------------------------------------
// The forward and inverse kernels load the forward weights of all L layers
// into LDS once (97.5 KB, one block per CU) and then stride over the samples
// with no further barrier: one 512-thread block (8 wavefronts, two per SIMD)
// takes 8 waves x TILES tiles x 16 samples per pass.  Tiles per wavefront: 4
// for long runs; batches below FWD_SMALL_N samples take FWD_TILES_SMALL so
// that the grid still covers every CU.
constexpr int FWD_WAVES = 8, FWD_TILES = 4, FWD_TILES_SMALL = 2, BWD_TILES = 4;
constexpr uint64_t FWD_SMALL_N = 1u << 20;

__device__ __forceinline__ void load_resident(LdsFwd (&s)[L], const uint16_t *wbf)
{
    for (int l = 0; l < L; ++l)
        load_layer<false>(s[l], wbf + (size_t)l * LP);
    __syncthreads();
}

template <int TILES>
__global__ __launch_bounds__(64 * FWD_WAVES) void flow_forward_kernel(FlowKernelArgs a)
{
    constexpr int CHUNK = FWD_WAVES * TILES * 16;
    __shared__ LdsFwd s[L];
    load_resident(s, a.wbf);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    for (uint64_t base = (uint64_t)blockIdx.x * CHUNK; base < a.n; base += (uint64_t)gridDim.x * CHUNK) {
        f4 z[TILES];
        float ld[TILES];
        bool valid[TILES];
        uint64_t idx[TILES];
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            idx[t] = base + (uint64_t)(wave * TILES + t) * 16 + (lane & 15);
            valid[t] = idx[t] < a.n;
            f4 vv = {0.5f, 0.5f, 0.5f, 0.5f};
            if (valid[t]) {
                if (a.v_in) {
                    vv = *reinterpret_cast<const f4 *>(a.v_in + 16 * idx[t] + 4 * g);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (4 * g + i < D)
                            vv[i] = stream_v(a.key, 12ull * (a.sample0 + idx[t]) + 4 * g + i);
                }
            }
            float l = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float x = 4 * g + i < D ? vv[i] : 0.5f;
                const float lx = flog(x), l1x = flog(1.f - x);
                z[t][i] = lx - l1x;
                if (4 * g + i < D)
                    l += lx + l1x; // log v (1 - v)
            }
            ld[t] = l;
        }
        for (int l = 0; l < L; ++l) {
            const uint32_t cond = COND_MASK[l];
#pragma unroll
            for (int t = 0; t < TILES; ++t) {
                if (a.zsave && valid[t])
                    *reinterpret_cast<f4 *>(a.zsave + ((size_t)l * a.n + idx[t]) * 16 + 4 * g) = z[t];
                MlpFwd r;
                mlp_forward(s[l], masked(z[t], cond), r);
#pragma unroll
                for (int i = 0; i < 4; ++i) { // branch-free: exp(0) = 1 exactly
                    const bool tr = transformed(cond, 4 * g + i);
                    const float sc = tr ? FLOW_SCALE * ftanh(r.o0[i]) : 0.f;
                    z[t][i] = z[t][i] * fexp(sc) + (tr ? r.o1[i] : 0.f);
                    ld[t] -= sc;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            float l = ld[t];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * g + i < D) // - log u (1 - u) = softplus(-z) + softplus(z)
                    l += softplus(-z[t][i]) + softplus(z[t][i]);
            l = sample_sum(l);
            if (!valid[t])
                continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * g + i;
                if (j < D) {
                    const double u = sigmoid_u(z[t][i]);
                    if (a.u_out)
                        a.u_out[a.u_stride ? j * a.u_stride + idx[t] : 12 * idx[t] + j] = u;
                    if (a.z_out)
                        a.z_out[12 * idx[t] + j] = z[t][i];
                }
            }
            if (g == 0) {
                if (a.logq_out)
                    a.logq_out[idx[t]] = l;
                if (a.invq_out)
                    a.invq_out[idx[t]] = 1. / (a.alpha * (double)fexp(l) + (1. - a.alpha));
            }
        }
    }
}

// ---- inverse: density of given or uniformly drawn points --------------------
template <int TILES>
__global__ __launch_bounds__(64 * FWD_WAVES) void flow_inverse_kernel(FlowKernelArgs a)
{
    constexpr int CHUNK = FWD_WAVES * TILES * 16;
    __shared__ LdsFwd s[L];
    load_resident(s, a.wbf);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    for (uint64_t base = (uint64_t)blockIdx.x * CHUNK; base < a.n; base += (uint64_t)gridDim.x * CHUNK) {
        f4 z[TILES];
        float ld[TILES];
        bool valid[TILES];
        uint64_t idx[TILES];
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            idx[t] = base + (uint64_t)(wave * TILES + t) * 16 + (lane & 15);
            valid[t] = idx[t] < a.n;
            float l = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * g + i;
                double u = 0.5;
                if (valid[t] && j < D) {
                    if (a.u_in) {
                        u = fmin(fmax(a.u_in[12 * idx[t] + j], 0x1.0p-126), 0x1.fffffffffffffp-1);
                    } else {
                        u = stream_u(a.key, 12ull * (a.sample0 + idx[t]) + j);
                        if (a.u_out)
                            a.u_out[a.u_stride ? j * a.u_stride + idx[t] : 12 * idx[t] + j] = u;
                    }
                }
                // u and 1 - u formed in f64, logs in f32
                const float lu = flog((float)u), l1u = flog((float)(1. - u));
                z[t][i] = lu - l1u;
                if (j < D)
                    l -= lu + l1u; // - log u (1 - u)
            }
            ld[t] = l;
        }
        for (int l = L - 1; l >= 0; --l) {
            const uint32_t cond = COND_MASK[l];
#pragma unroll
            for (int t = 0; t < TILES; ++t) {
                MlpFwd r;
                mlp_forward(s[l], masked(z[t], cond), r);
#pragma unroll
                for (int i = 0; i < 4; ++i) { // branch-free: exp(0) = 1 exactly
                    const bool tr = transformed(cond, 4 * g + i);
                    const float sc = tr ? FLOW_SCALE * ftanh(r.o0[i]) : 0.f;
                    z[t][i] = (z[t][i] - (tr ? r.o1[i] : 0.f)) * fexp(-sc);
                    ld[t] -= sc;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            float l = ld[t];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * g + i < D) // log v (1 - v) = -softplus(-z) - softplus(z)
                    l -= softplus(-z[t][i]) + softplus(z[t][i]);
            l = sample_sum(l);
            if (valid[t] && g == 0) {
                if (a.logq_out)
                    a.logq_out[idx[t]] = l;
                if (a.invq_out)
                    a.invq_out[idx[t]] = 1. / (a.alpha * (double)fexp(l) + (1. - a.alpha));
            }
        }
    }
}

// ---- backward (training) -----------------------------------------------------
// loss per sample: log q(u) - log(f(u) + eps); gz = dloss/dz at the flow
// output.  Walks the layers top-down, recomputing each MLP from the saved
// layer input, and stores the bf16 activations and deltas (feature-major) for
// the batch weight-gradient products.
template <int TILES>
__global__ __launch_bounds__(256) void flow_backward_kernel(FlowBackArgs a)
{
    constexpr int CHUNK = 4 * TILES * 16;
    __shared__ LdsLayer s;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    for (uint64_t base = (uint64_t)blockIdx.x * CHUNK; base < a.n; base += (uint64_t)gridDim.x * CHUNK) {
        f4 gz[TILES];
        bool valid[TILES];
        uint64_t idx[TILES];
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            idx[t] = base + (uint64_t)(wave * TILES + t) * 16 + (lane & 15);
            valid[t] = idx[t] < a.n;
            gz[t] = valid[t] ? *reinterpret_cast<const f4 *>(a.gz + 16 * idx[t] + 4 * g) : (f4){0.f, 0.f, 0.f, 0.f};
        }
        for (int l = L - 1; l >= 0; --l) {
            __syncthreads();
            load_layer<true>(s, a.wbf + (size_t)l * LP);
            __syncthreads();
            const uint32_t cond = COND_MASK[l];
            uint16_t *sv = a.save + (size_t)l * SAVE_F * a.n;
#pragma unroll 1
            for (int t = 0; t < TILES; ++t) {
                f4 zin = valid[t] ? *reinterpret_cast<const f4 *>(a.zsave + ((size_t)l * a.n + idx[t]) * 16 + 4 * g)
                                  : (f4){0.f, 0.f, 0.f, 0.f};
                // rows row0 + 4g .. + 3 of this lane's sample, stored as soon as
                // they are formed (short live ranges)
                const bool ok = valid[t];
                const uint64_t n = idx[t], N = a.n;
                auto put = [&](int row0, s4 v) {
                    if (ok)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            sv[(size_t)(row0 + 4 * g + i) * N + n] = (uint16_t)v[i];
                };
                MlpFwd r;
                mlp_forward(s, masked(zin, cond), r);
                put(SAVE_A, r.a);
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) {
                    put(SAVE_H1 + 16 * mb, r.h1[mb]);
                    put(SAVE_H2 + 16 * mb, r.h2[mb]);
                }
                f4 ob0, ob1, zb;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int j = 4 * g + i;
                    if (transformed(cond, j)) {
                        const float th = ftanh(r.o0[i]);
                        const float es = fexp(FLOW_SCALE * th);
                        const float sbar = gz[t][i] * zin[i] * es - 1.f; // the -sum s term of log q
                        ob0[i] = sbar * FLOW_SCALE * (1.f - th * th);
                        ob1[i] = gz[t][i];
                        zb[i] = gz[t][i] * es;
                    } else {
                        ob0[i] = 0.f;
                        ob1[i] = 0.f;
                        zb[i] = j < D ? gz[t][i] : 0.f;
                    }
                }
                const s4 obb0 = pack4(ob0), obb1 = pack4(ob1);
                put(SAVE_OB, obb0);
                put(SAVE_OB + 16, obb1);
                const f4 zero = {0.f, 0.f, 0.f, 0.f};
                s4 h2b[4], h1b[4];
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) {
                    f4 c = mfma(atile(s.w + L_W3T, S_W3T, mb, 0), obb0, zero);
                    c = mfma(atile(s.w + L_W3T, S_W3T, mb, 1), obb1, c);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        c[i] *= r.h2d[mb][i];
                    h2b[mb] = pack4(c);
                    put(SAVE_H2B + 16 * mb, h2b[mb]);
                }
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) {
                    f4 c = zero;
#pragma unroll
                    for (int kb = 0; kb < 4; ++kb)
                        c = mfma(atile(s.w + L_W2T, S_W2T, mb, kb), h2b[kb], c);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        c[i] *= r.h1d[mb][i];
                    h1b[mb] = pack4(c);
                    put(SAVE_H1B + 16 * mb, h1b[mb]);
                }
                f4 ab = zero;
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    ab = mfma(atile(s.w + L_W1T, S_W1T, 0, kb), h1b[kb], ab);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (conditioning(cond, 4 * g + i))
                        zb[i] += ab[i];
                gz[t] = zb;
            }
        }
        if (a.gv)
#pragma unroll
            for (int t = 0; t < TILES; ++t)
                if (valid[t])
                    *reinterpret_cast<f4 *>(a.gv + 16 * idx[t] + 4 * g) = gz[t];
    }
}

// ---- batch weight and bias gradients -----------------------------------------
// dW[m][k] = inv_n sum_n delta[m][n] act[k][n] and db[m] = inv_n sum_n delta[m][n]
// on MFMA, with the sample index as the reduction dimension.  One 256-thread
// block per (layer, sample range): chunks of 64 samples of all 304 saved rows
// are staged through LDS (the next chunk is loaded into registers while the
// current one is multiplied), so every saved value is read from HBM once.
// The layer's 38 output tiles -- 28 weight tiles and 10 bias tiles, whose
// B operand is a column of ones -- are spread over the four wavefronts.  Each
// block writes partial sums for all of its layer's parameters, and
// flow_grad_reduce_kernel adds the partials of each parameter in block order,
// so the result does not depend on scheduling.
constexpr int GR_CHUNK = 64, GR_STRIDE = 68, GR_TILES = 38, GR_PER_WAVE = 10;
constexpr int GR_PIECES = SAVE_F * (GR_CHUNK / 8), GR_LOADS = (GR_PIECES + 255) / 256;
constexpr int GR_MIN_CHUNKS = 8, GR_MAX_SPLIT = 128;

struct GTile {
    int delta_row, act_row, param_off, ld; // act_row < 0: bias tile (ones operand)
};

// Tiles of one layer: W1 4 row blocks, W2 4 x 4, W3 2 x 4, then b1 4, b2 4,
// b3 2 row blocks.  param_off is the parameter of the tile's element (0, 0).
__device__ __forceinline__ GTile gtile(int t)
{
    if (t < 4)
        return {SAVE_H1B + 16 * t, SAVE_A, OFF_W1 + 16 * t * DP, DP};
    if (t < 20) {
        const int mb = (t - 4) / 4, kb = (t - 4) % 4;
        return {SAVE_H2B + 16 * mb, SAVE_H1 + 16 * kb, OFF_W2 + 16 * mb * H + 16 * kb, H};
    }
    if (t < 28) {
        const int mb = (t - 20) / 4, kb = (t - 20) % 4;
        return {SAVE_OB + 16 * mb, SAVE_H2 + 16 * kb, OFF_W3 + 16 * mb * H + 16 * kb, H};
    }
    if (t < 32)
        return {SAVE_H1B + 16 * (t - 28), -1, OFF_B1 + 16 * (t - 28), 0};
    if (t < 36)
        return {SAVE_H2B + 16 * (t - 32), -1, OFF_B2 + 16 * (t - 32), 0};
    return {SAVE_OB + 16 * (t - 36), -1, OFF_B3 + 16 * (t - 36), 0};
}

// Sample-range blocks per layer for n samples.
static int wgrad_split(uint64_t n)
{
    const uint64_t chunks = (n + GR_CHUNK - 1) / GR_CHUNK;
    const uint64_t s = (chunks + GR_MIN_CHUNKS - 1) / GR_MIN_CHUNKS;
    return (int)(s < 1 ? 1 : s > GR_MAX_SPLIT ? GR_MAX_SPLIT : s);
}

__global__ __launch_bounds__(256) void flow_wgrad_kernel(const uint16_t *save, uint64_t n, int split, float *part)
{
    __shared__ uint16_t lds[SAVE_F * GR_STRIDE];
    const int layer = blockIdx.x / split, ks = blockIdx.x % split;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, r16 = lane & 15;
    const uint64_t chunks = (n + GR_CHUNK - 1) / GR_CHUNK, per = (chunks + split - 1) / split;
    const uint64_t c0 = (uint64_t)ks * per, c1 = c0 + per < chunks ? c0 + per : chunks;
    const uint16_t *sv = save + (size_t)layer * SAVE_F * n;
    // piece p = 8 samples (16 B) of row p / 8; n is a multiple of 16, so a
    // piece lies wholly inside or wholly past the batch
    uint4 buf[GR_LOADS];
    auto load = [&](uint64_t c) {
#pragma unroll
        for (int k = 0; k < GR_LOADS; ++k) {
            const int p = threadIdx.x + 256 * k;
            const uint64_t smp = c * GR_CHUNK + 8 * (p & 7);
            buf[k] = make_uint4(0u, 0u, 0u, 0u);
            if (p < GR_PIECES && smp < n)
                buf[k] = *reinterpret_cast<const uint4 *>(sv + (size_t)(p >> 3) * n + smp);
        }
    };
    f4 acc[GR_PER_WAVE];
#pragma unroll
    for (int j = 0; j < GR_PER_WAVE; ++j)
        acc[j] = (f4){0.f, 0.f, 0.f, 0.f};
    const short one = (short)0x3F80; // bf16 1.0
    const s4 ones = {one, one, one, one};
    if (c0 < c1)
        load(c0);
    for (uint64_t c = c0; c < c1; ++c) {
        __syncthreads(); // the previous chunk's products are done
#pragma unroll
        for (int k = 0; k < GR_LOADS; ++k) {
            const int p = threadIdx.x + 256 * k;
            if (p < GR_PIECES) {
                uint2 *d = reinterpret_cast<uint2 *>(lds + (p >> 3) * GR_STRIDE + 8 * (p & 7));
                d[0] = make_uint2(buf[k].x, buf[k].y);
                d[1] = make_uint2(buf[k].z, buf[k].w);
            }
        }
        __syncthreads();
        if (c + 1 < c1)
            load(c + 1);
#pragma unroll
        for (int kk = 0; kk < GR_CHUNK / 16; ++kk)
#pragma unroll
            for (int j = 0; j < GR_PER_WAVE; ++j) {
                const int t = wave + 4 * j;
                if (t < GR_TILES) {
                    const GTile w = gtile(t);
                    const int col = 16 * kk + 4 * g;
                    const s4 A = *reinterpret_cast<const s4 *>(lds + (w.delta_row + r16) * GR_STRIDE + col);
                    const s4 B =
                        w.act_row < 0 ? ones : *reinterpret_cast<const s4 *>(lds + (w.act_row + r16) * GR_STRIDE + col);
                    acc[j] = mfma(A, B, acc[j]);
                }
            }
    }
    float *pl = part + ((size_t)ks * L + layer) * LP;
#pragma unroll
    for (int j = 0; j < GR_PER_WAVE; ++j) {
        const int t = wave + 4 * j;
        if (t >= GR_TILES)
            continue;
        const GTile w = gtile(t);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i; // C layout: lane holds rows 4g.. of column r16
            if (w.act_row >= 0)
                pl[w.param_off + row * w.ld + r16] = acc[j][i];
            else if (r16 == 0)
                pl[w.param_off + row] = acc[j][i];
        }
    }
}

// grad[i] = inv_n sum over the sample-range blocks, in block order (loads
// in groups of 16 so that they are in flight together).
__global__ __launch_bounds__(256) void flow_grad_reduce_kernel(const float *part, int split, float inv_n, float *grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM)
        return;
    float sum = 0.f;
    for (int k0 = 0; k0 < split; k0 += 16) {
        float v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            v[j] = k0 + j < split ? part[(size_t)(k0 + j) * NPARAM + i] : 0.f;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            sum += v[j];
    }
    grad[i] = sum * inv_n;
}

__global__ void flow_adam_kernel(float *w, uint16_t *wbf, float *m, float *v, const float *g, int step, float lr,
                                 float b1, float b2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)step)), vh = vi / (1.f - powf(b2, (float)step));
    const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    wbf[i] = f2bf(wi);
}

__global__ void flow_to_bf16_kernel(const float *w, uint16_t *wbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NPARAM)
        wbf[i] = f2bf(w[i]);
}

// ---- launchers ---------------------------------------------------------------
template <int TILES> static unsigned grid_for(uint64_t n) { return (unsigned)((n + 64 * TILES - 1) / (64 * TILES)); }

// Resident-weight kernels: one pass per block covers FWD_WAVES * TILES * 16
// samples; at most max_blocks blocks (one per CU), striding over the rest.
template <int TILES> static unsigned grid_resident(const FlowKernelArgs &a)
{
    const uint64_t chunk = (uint64_t)FWD_WAVES * TILES * 16, need = (a.n + chunk - 1) / chunk;
    const uint64_t cap = a.max_blocks ? a.max_blocks : need;
    return (unsigned)std::max<uint64_t>(1, std::min(need, cap));
}

void launch_flow_forward(const FlowKernelArgs &a, hipStream_t st)
{
    if (a.n < FWD_SMALL_N)
        hipLaunchKernelGGL(flow_forward_kernel<FWD_TILES_SMALL>, dim3(grid_resident<FWD_TILES_SMALL>(a)),
                           dim3(64 * FWD_WAVES), 0, st, a);
    else
        hipLaunchKernelGGL(flow_forward_kernel<FWD_TILES>, dim3(grid_resident<FWD_TILES>(a)), dim3(64 * FWD_WAVES), 0,
                           st, a);
}

void launch_flow_inverse(const FlowKernelArgs &a, hipStream_t st)
{
    if (a.n < FWD_SMALL_N)
        hipLaunchKernelGGL(flow_inverse_kernel<FWD_TILES_SMALL>, dim3(grid_resident<FWD_TILES_SMALL>(a)),
                           dim3(64 * FWD_WAVES), 0, st, a);
    else
        hipLaunchKernelGGL(flow_inverse_kernel<FWD_TILES>, dim3(grid_resident<FWD_TILES>(a)), dim3(64 * FWD_WAVES), 0,
                           st, a);
}

void launch_flow_backward(const FlowBackArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(flow_backward_kernel<BWD_TILES>, dim3(grid_for<BWD_TILES>(a.n)), dim3(256), 0, st, a);
}

size_t flow_wgrad_part_bytes(uint64_t n) { return (size_t)wgrad_split(n) * NPARAM * sizeof(float); }

void launch_flow_wgrad(const uint16_t *save, uint64_t n, float inv_n, float *part, float *grad, hipStream_t st)
{
    const int split = wgrad_split(n);
    hipLaunchKernelGGL(flow_wgrad_kernel, dim3(L * split), dim3(256), 0, st, save, n, split, part);
    hipLaunchKernelGGL(flow_grad_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, part, split, inv_n,
                       grad);
}

void launch_flow_adam(float *w, uint16_t *wbf, float *m, float *v, const float *g, int step, float lr, float beta1,
                      float beta2, float eps, hipStream_t st)
{
    hipLaunchKernelGGL(flow_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, w, wbf, m, v, g, step, lr,
                       beta1, beta2, eps);
}

void launch_flow_to_bf16(const float *w, uint16_t *wbf, hipStream_t st)
{
    hipLaunchKernelGGL(flow_to_bf16_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, w, wbf);
}

// ---- integrand and its gradient for training --------------------------------
// Physics from here on: no contraction, as in the rest of the build.
#pragma clang fp contract(off)

// f(u) = sum_c lam_c |M|^2_c(u) cut(u) from the hand-written adjoint (the
// smooth cut factor of event_adjoint when tau > 0), and
// gz = dloss/dz at the flow output for loss = log q - log(f + eps):
//   dloss/du = -(1 - 2u)/(u (1 - u)) - f'(u)/(f + eps),  dz = du * u (1 - u).
__global__ __launch_bounds__(128) void flow_integrand_kernel(tpd::Params P, double lam0, double lam1, double lam2,
                                                             double lam3, double lam4, uint64_t n, const double *u,
                                                             double eps, double tau, double fcap, int target,
                                                             float clip, float *gz, float *logf, double *f_out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const tpd::AdjParams<double> A = tpd::adj_params<double>(P, 1.0);
    double uu[12], m[5], du[12], dp[12];
    const double lam[5] = {lam0, lam1, lam2, lam3, lam4};
#pragma unroll
    for (int j = 0; j < 12; ++j)
        uu[j] = u[12 * i + j];
    bool keep;
    if (target == 2)
        keep = tpd::event_adjoint<double, 2>(A, uu, lam, m, du, dp, tau, fcap);
    else if (target == 1)
        keep = tpd::event_adjoint<double, 1>(A, uu, lam, m, du, dp, tau);
    else
        keep = tpd::event_adjoint<double, 0>(A, uu, lam, m, du, dp);
    double f = 0.;
    if (keep)
        for (int c = 0; c < 5; ++c)
            f += lam[c] * m[c];
    // target 2 differentiates log f itself (du already carries the 1/f)
    const double fe = target == 2 ? (keep ? 1. : INFINITY) : f + eps;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        const double w = uu[j] * (1. - uu[j]);
        const float g = (float)(-(1. - 2. * uu[j]) - w * du[j] / fe);
        gz[16 * i + j] = fminf(fmaxf(g, -clip), clip);
    }
#pragma unroll
    for (int j = 12; j < 16; ++j)
        gz[16 * i + j] = 0.f;
    logf[i] = (float)(target == 2 ? (keep ? log(f) : -100.) : log(fe)); // dropped events: -100
    if (f_out)
        f_out[i] = f;
}

void launch_flow_integrand(const tpd::Params &P, const double lam[5], uint64_t n, const double *u, double eps,
                           double tau, double fcap, int target, float clip, float *gz, float *logf, double *f_out,
                           hipStream_t st)
{
    hipLaunchKernelGGL(flow_integrand_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, P, lam[0], lam[1],
                       lam[2], lam[3], lam[4], n, u, eps, tau, fcap, target, clip, gz, logf, f_out);
}

// Mean over the batch of log q - log(f + eps), one 1024-thread block, fixed
// summation order (per-thread strided sums of 4-wide loads, then a tree).
__global__ __launch_bounds__(1024) void flow_loss_kernel(const float *logq, const float *logf, uint64_t n, float *out)
{
    __shared__ double red[1024];
    double acc = 0.;
    const uint64_t n4 = n / 4;
    const float4 *q4 = reinterpret_cast<const float4 *>(logq), *f4p = reinterpret_cast<const float4 *>(logf);
#pragma unroll 4
    for (uint64_t i = threadIdx.x; i < n4; i += 1024) {
        const float4 q = q4[i], f = f4p[i];
        acc += (((double)q.x - (double)f.x) + ((double)q.y - (double)f.y)) +
               (((double)q.z - (double)f.z) + ((double)q.w - (double)f.w));
    }
    for (uint64_t i = 4 * n4 + threadIdx.x; i < n; i += 1024)
        acc += (double)logq[i] - (double)logf[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = (float)(red[0] / (double)n);
}

void launch_flow_loss(const float *logq, const float *logf, uint64_t n, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(flow_loss_kernel, dim3(1), dim3(1024), 0, st, logq, logf, n, out);
}

} // namespace tpf
