// tp_flow.h -- the learned proposal: a RealNVP flow on the 12 RAMBO uniforms.
//
// v ~ U(0,1)^12 (from the xoshiro stream) -> z = logit(v) -> six masked affine
// coupling layers -> u = sigmoid(z).  Coupling layer l keeps the features
// in cond_mask[l] and maps the others as z_j <- z_j exp(s_j) + t_j with
// (s, t) = MLP_l(z * cond_mask[l]), MLP = 12 -> 64 -> 64 -> 24, hard-swish, and
// s = FLOW_SCALE * tanh(raw).  The density of u is
//   log q(u) = sum_j log(v_j (1 - v_j)) - sum_l sum_j s_lj - sum_j log(u_j (1 - u_j)).
// The three matrix products of every MLP run on bf16 MFMA with f32
// accumulation; everything else is f32.
//
// The flow's own randomness (v, and the uniform points of the mixture) is the
// SplitMix64 sequence used as a counter-based generator: draw k of key is
// mix64(key + (k + 1) * 0x9E3779B97F4A7C15) (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014), so every lane
// computes its own draws with no skip-ahead and no buffer.
//
// Parameter block of one layer (f32 master copy and bf16 copy, same layout):
//   W1[64][16] (input features 12..15 are zero padding), b1[64],
//   W2[64][64], b2[64], W3[32][64], b3[32] with rows 0..11 = s outputs and
//   rows 16..27 = t outputs (rows 12..15, 28..31 zero).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpf {

constexpr int D = 12, DP = 16, H = 64, O = 32, L = 6;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * DP, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + O * H, LP = OFF_B3 + O; // 7328 per layer
constexpr int NPARAM = L * LP;
constexpr float FLOW_SCALE = 1.5f;
// Features kept (conditioning) by each layer; each feature is transformed by
// three of the six layers.
constexpr uint32_t COND_MASK[L] = {0x03Fu, 0xFC0u, 0x555u, 0xAAAu, 0x1C7u, 0xE38u};
// Values of the batch activations saved for the weight gradients, per layer
// and sample: a(16) h1(64) h2(64) obar(32) h2bar(64) h1bar(64).
constexpr int SAVE_A = 0, SAVE_H1 = 16, SAVE_H2 = 80, SAVE_OB = 144, SAVE_H2B = 176, SAVE_H1B = 240, SAVE_F = 304;

struct FlowKernelArgs {
    const uint16_t *wbf; // bf16 parameters [L][LP]
    uint64_t n;          // samples
    // Input of the forward pass: explicit v (n x 16 f32, device) when v_in is
    // set, else the counter stream: v_ij from draw 12 (sample0 + i) + j of key.
    const float *v_in;
    uint64_t key;
    uint64_t sample0;    // stream offset of sample 0, in samples
    // outputs (each may be null)
    double *u_out;       // [n][12] f64, or [12][u_stride] when u_stride > 0
    uint64_t u_stride;
    float *z_out;        // [n][12] f32 flow output z = logit(u)
    float *logq_out;     // [n]
    float *zsave;        // [L][n][16] layer inputs (training)
    // Input of the inverse pass: u_in [n][12] f64 when set, else uniform u from
    // the counter stream (written to u_out).
    const double *u_in;
    // mixture bookkeeping for the weighted event loop: q_mix = alpha q + (1 - alpha)
    double alpha;
    double *invq_out;    // [n] 1 / q_mix
    uint32_t max_blocks; // grid cap of the forward and inverse kernels (the CU count), 0: none
};

struct FlowBackArgs {
    const uint16_t *wbf;
    uint64_t n;
    const float *zsave;  // [L][n][16]
    const float *gz;     // [n][16] dloss/dz at the flow output
    uint16_t *save;      // [L][SAVE_F][n] bf16, feature-major
    float *gv;           // [n][16] dloss/dz at the flow input (may be null)
};

// u = T(v), log q (and 1/q_mix, z per layer).
void launch_flow_forward(const FlowKernelArgs &a, hipStream_t st);
// log q (and 1/q_mix) of given or uniformly drawn points.
void launch_flow_inverse(const FlowKernelArgs &a, hipStream_t st);
void launch_flow_backward(const FlowBackArgs &a, hipStream_t st);
// Weight and bias gradients over the batch (n a multiple of 16):
// grad[L][LP] = inv_n sum_n ..., through a partial-sum buffer `part` of
// flow_wgrad_part_bytes(n) bytes.
size_t flow_wgrad_part_bytes(uint64_t n);
void launch_flow_wgrad(const uint16_t *save, uint64_t n, float inv_n, float *part, float *grad, hipStream_t st);
void launch_flow_adam(float *w, uint16_t *wbf, float *m, float *v, const float *g, int step, float lr, float beta1,
                      float beta2, float eps, hipStream_t st);
void launch_flow_to_bf16(const float *w, uint16_t *wbf, hipStream_t st);

} // namespace tpf

namespace tpd {
struct Params;
}
namespace tpf {
// Training integrand and the gradient of the loss log q - log g at the flow
// output z (gz, each component clipped to [-clip, clip]), with f = sum_c
// lam_c |M|^2_c (RAMBO weight 1) and S the cuts smoothed over tau:
//   target 0: g = f cut + eps;  target 1: g = f S + eps;
//   target 2: g = min(f, fcap) S (log g = log min(f, fcap) + log S,
//   differentiated directly).
// f_out = f cut (0) or g (1, 2); logf = log g (-100 for events target 2 drops).
void launch_flow_integrand(const tpd::Params &P, const double lam[5], uint64_t n, const double *u, double eps,
                           double tau, double fcap, int target, float clip, float *gz, float *logf, double *f_out,
                           hipStream_t st);
// *out = mean(logq - logf) over n samples.
void launch_flow_loss(const float *logq, const float *logf, uint64_t n, float *out, hipStream_t st);
} // namespace tpf
