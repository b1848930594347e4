// Learned importance-sampling proposal: a coupling-layer normalising flow with
// rational-quadratic spline bijectors (Durkan, Bekasov, Murray, Papamakarios,
// arXiv:1906.04032) on the VEGAS-mapped unit hypercube.  It has no counterpart in the
// reference; it sits between the uniforms and the VEGAS map of the integration loop.
//
// Layer l splits the dimensions into a conditioning set A_l and a transformed set B_l.
// An MLP  x_A -> 128 -> 128 -> |B_l| (3K+1)  (ReLU) gives, per transformed dimension,
// K bin widths, K bin heights (softmax, minimum 1e-3) and K+1 knot derivatives
// (softplus + 1e-3) of a monotone RQ spline on [0,1].
//   forward  x -> u  with log|det du/dx|   (density: log q(x) = sum of log-dets)
//   inverse  u -> x  (sampling; analytic root of the RQ quadratic)
//   backward gradients of a per-sample weighted objective through every layer
// Kernels work on one block of 8 waves per 16 samples (the waves split the MFMA output
// tiles, so that a small training batch still puts two waves on every SIMD): the 128x128
// and 128xNP GEMMs of the forward
// pass run on bf16 MFMA (v_mfma_f32_16x16x16bf16, fp32 accumulation, bf16 weights),
// the backward GEMMs on fp32 MFMA (v_mfma_f32_16x16x4f32); the spline math is per-lane
// VALU.  Weight-gradient partials are written per block and reduced in a fixed order.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "mfx_api_common.h"
#include "../../include/madflow_amd.h"

namespace mfx {
namespace flow {

constexpr int H = 128;          // hidden width
constexpr int K = 16;           // spline bins
constexpr int NPD = 3 * K + 1;  // raw spline parameters per transformed dimension
constexpr int DMAX = 16;        // maximum flow dimension
constexpr int TILE = 16;        // samples per block
constexpr int WPB = 8;          // waves per block
constexpr int NT = 64 * WPB;    // threads per block
constexpr float MINW = 1e-3f, MINH = 1e-3f, MIND = 1e-3f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

struct LayerGeom {
  int nA, nB, NP;                  // NP = padded |B| * NPD (multiple of 16)
  int A[DMAX], B[DMAX];
  // offsets (in floats) of this layer's parameters in the fp32 parameter vector
  int oW1, ob1, oW2, ob2, oW3, ob3;
  // offsets (in bf16 elements) of the bf16 copies: transposed W2t [H][H], W3t [NP][H]
  // (B operands of the forward GEMMs) and in the original layout W2n [H][H], W3n [H][NP]
  // (B operands of the backward GEMMs)
  int oW2t, oW3t, oW2n, oW3n;
};

struct FlowGeom {
  int D, L, NPmax;
  int nparams, nbf16;
  LayerGeom layer[16];
};

// ------------------------------------------------------------------ spline math
struct Knots {
  float X, w, Y, h, a, c;   // bin left edge, width, height offset, height, derivatives
  int k;
};

// widths/heights from raw logits, softmax with a minimum; returns the cumulative edge
// below ``v`` (searching in the cumulative sums) and the bin.
__device__ __forceinline__ void softmax_bins(const float* r, float minv, float* out) {
  float m = r[0];
  for (int j = 1; j < K; ++j) m = fmaxf(m, r[j]);
  float s = 0.f;
  for (int j = 0; j < K; ++j) {
    out[j] = __expf(r[j] - m);
    s += out[j];
  }
  const float sc = (1.f - K * minv) / s;
  for (int j = 0; j < K; ++j) out[j] = minv + sc * out[j];
}
__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(__expf(x)); }
__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + __expf(-x)); }

// Locate the bin of ``v`` along the cumulative sums of ``wdt`` (edge search variable:
// x for the forward map, the heights for the inverse).
__device__ __forceinline__ int find_bin(const float* cum, float v) {
  int k = 0;
  for (int j = 1; j < K; ++j) k += (v >= cum[j]) ? 1 : 0;
  return k;
}

struct SplineEval {
  float z, ld;  // output and log derivative
};

__device__ __forceinline__ void knots_of(const float* raw, float* w, float* h, float* cw, float* ch) {
  softmax_bins(raw, MINW, w);
  softmax_bins(raw + K, MINH, h);
  cw[0] = 0.f;
  ch[0] = 0.f;
  for (int j = 0; j < K; ++j) {
    cw[j + 1] = cw[j] + w[j];
    ch[j + 1] = ch[j] + h[j];
  }
  cw[K] = 1.f;
  ch[K] = 1.f;
}

__device__ __forceinline__ SplineEval rq_forward(const float* raw, float x) {
  float w[K], h[K], cw[K + 1], ch[K + 1];
  knots_of(raw, w, h, cw, ch);
  x = fminf(fmaxf(x, 0.f), 1.f);
  const int k = find_bin(cw, x);
  const float a = MIND + softplus(raw[2 * K + k]), c = MIND + softplus(raw[2 * K + k + 1]);
  const float wk = cw[k + 1] - cw[k], hk = ch[k + 1] - ch[k];
  const float xi = fminf(fmaxf((x - cw[k]) / wk, 0.f), 1.f);
  const float s = hk / wk;
  const float t = xi * (1.f - xi);
  const float num = hk * (s * xi * xi + a * t);
  const float den = s + (a + c - 2.f * s) * t;
  const float P = c * xi * xi + 2.f * s * t + a * (1.f - xi) * (1.f - xi);
  SplineEval r;
  r.z = ch[k] + num / den;
  r.ld = 2.f * __logf(s) + __logf(P) - 2.f * __logf(den);
  return r;
}

__device__ __forceinline__ SplineEval rq_inverse(const float* raw, float z) {
  float w[K], h[K], cw[K + 1], ch[K + 1];
  knots_of(raw, w, h, cw, ch);
  z = fminf(fmaxf(z, 0.f), 1.f);
  const int k = find_bin(ch, z);
  const float a = MIND + softplus(raw[2 * K + k]), c = MIND + softplus(raw[2 * K + k + 1]);
  const float wk = cw[k + 1] - cw[k], hk = ch[k + 1] - ch[k];
  const float s = hk / wk;
  const float d = z - ch[k];
  const float g = a + c - 2.f * s;
  const float qa = hk * (s - a) + d * g;
  const float qb = hk * a - d * g;
  const float qc = -s * d;
  const float disc = fmaxf(qb * qb - 4.f * qa * qc, 0.f);
  float xi = (2.f * qc) / (-qb - sqrtf(disc));
  xi = fminf(fmaxf(xi, 0.f), 1.f);
  const float t = xi * (1.f - xi);
  const float den = s + g * t;
  const float P = c * xi * xi + 2.f * s * t + a * (1.f - xi) * (1.f - xi);
  SplineEval r;
  r.z = cw[k] + xi * wk;                                          // x
  r.ld = -(2.f * __logf(s) + __logf(P) - 2.f * __logf(den));     // log|dx/du|
  return r;
}

// Gradient of gz*z + gl*ld w.r.t. the raw parameters (written to graw[NPD]) and x.
__device__ __forceinline__ float rq_backward(const float* raw, float x, float gz, float gl, float* graw) {
  float w[K], h[K], cw[K + 1], ch[K + 1];
  knots_of(raw, w, h, cw, ch);
  const bool inside = x > 0.f && x < 1.f;
  x = fminf(fmaxf(x, 0.f), 1.f);
  const int k = find_bin(cw, x);
  const float ra = raw[2 * K + k], rc = raw[2 * K + k + 1];
  const float a = MIND + softplus(ra), c = MIND + softplus(rc);
  const float wk = cw[k + 1] - cw[k], hk = ch[k + 1] - ch[k];
  const float xi = fminf(fmaxf((x - cw[k]) / wk, 0.f), 1.f);
  const float s = hk / wk;
  const float t = xi * (1.f - xi);
  const float g2 = a + c - 2.f * s;
  const float N = hk * (s * xi * xi + a * t);
  const float Dn = s + g2 * t;
  const float P = c * xi * xi + 2.f * s * t + a * (1.f - xi) * (1.f - xi);
  const float iD = 1.f / Dn, iD2 = iD * iD, iP = 1.f / P;
  const float dt = 1.f - 2.f * xi;
  // partials of N, Dn, P
  const float Nxi = hk * (2.f * s * xi + a * dt), Ns = hk * xi * xi, Nh = s * xi * xi + a * t, Na = hk * t;
  const float Dxi = g2 * dt, Ds = 1.f - 2.f * t, Da = t, Dc = t;
  const float Pxi = 2.f * c * xi + 2.f * s * dt - 2.f * a * (1.f - xi), Ps = 2.f * t, Pa = (1.f - xi) * (1.f - xi),
              Pc = xi * xi;
  // combined gradient of gz*z + gl*ld w.r.t. (xi, s, h, Y, a, c)
  const float Gxi = gz * (Nxi * Dn - N * Dxi) * iD2 + gl * (Pxi * iP - 2.f * Dxi * iD);
  const float Gs = gz * (Ns * Dn - N * Ds) * iD2 + gl * (2.f / s + Ps * iP - 2.f * Ds * iD);
  const float Gh = gz * Nh * iD;
  const float GY = gz;
  const float Ga = gz * (Na * Dn - N * Da) * iD2 + gl * (Pa * iP - 2.f * Da * iD);
  const float Gc = gz * (-N * Dc) * iD2 + gl * (Pc * iP - 2.f * Dc * iD);
  // chain to bin quantities
  const float GX = -Gxi / wk;
  const float Gw = -Gxi * xi / wk - Gs * s / wk;
  const float Ghk = Gh + Gs / wk;
  float gw[K], gh[K];
  for (int j = 0; j < K; ++j) {
    gw[j] = (j < k ? GX : 0.f) + (j == k ? Gw : 0.f);
    gh[j] = (j < k ? GY : 0.f) + (j == k ? Ghk : 0.f);
  }
  // through the softmax: w = min + (1 - K min) sm
  float sw = 0.f, sh = 0.f;
  for (int j = 0; j < K; ++j) {
    sw += (w[j] - MINW) * gw[j];
    sh += (h[j] - MINH) * gh[j];
  }
  const float nw = 1.f - K * MINW, nh = 1.f - K * MINH;
  for (int j = 0; j < K; ++j) {
    graw[j] = (w[j] - MINW) * (gw[j] - sw / nw);
    graw[K + j] = (h[j] - MINH) * (gh[j] - sh / nh);
  }
  for (int j = 0; j <= K; ++j) graw[2 * K + j] = 0.f;
  graw[2 * K + k] = Ga * sigmoid(ra);
  graw[2 * K + k + 1] += Gc * sigmoid(rc);
  return inside ? Gxi / wk : 0.f;
}

// ------------------------------------------------------------------ MFMA helpers
// v_mfma_f32_16x16x16bf16_1k: lane l supplies A[l%16][4(l/16)..+3] and B[4(l/16)..+3][l%16];
// the 16x16 result has D[4(l/16)+i][l%16] in element i of lane l.
__device__ __forceinline__ f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// v_mfma_f32_16x16x4f32: lane l supplies A[l%16][l/16] and B[l/16][l%16].
__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ MLP forward
// One block, 16 samples; wave w computes the output tiles n0 = 16 (w + 4 i).
// xa: LDS [16][DMAX] conditioning inputs (only nA used); h1s/h2s: LDS bf16 [16][H];
// raws: LDS fp32 [16][NP].
// NPC: the layer's NP when known at compile time (0: g.NP at run time).
template <int NPC>
__device__ void mlp_forward(const LayerGeom& g, const float* __restrict__ P, const unsigned short* __restrict__ Pb,
                            const float* xa, unsigned short* h1s, unsigned short* h2s, float* raws) {
  const int NP = NPC ? NPC : g.NP;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // layer 1 on VALU: 16 x 128 outputs, 8 per thread
  for (int q = tid; q < TILE * H; q += NT) {
    const int s = q / H, j = q % H;
    float v = P[g.ob1 + j];
    for (int i = 0; i < g.nA; ++i) v += xa[s * DMAX + i] * P[g.oW1 + i * H + j];
    h1s[s * H + j] = f2bf(fmaxf(v, 0.f));
  }
  __syncthreads();
  const int r = lane & 15, kq = (lane >> 4) * 4;
  // layer 2: h2 = relu(h1 W2 + b2), B operand from W2t [out][in] (bf16)
  for (int n0 = 16 * wid; n0 < H; n0 += 16 * WPB) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < H; k0 += 16) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(h1s + r * H + k0 + kq);
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(Pb + g.oW2t + (n0 + r) * H + k0 + kq);
      acc = mfma_bf16(a, b, acc);
    }
    for (int i = 0; i < 4; ++i) {
      const int row = kq + i, col = n0 + r;
      h2s[row * H + col] = f2bf(fmaxf(acc[i] + P[g.ob2 + col], 0.f));
    }
  }
  __syncthreads();
  // layer 3: raw = h2 W3 + b3, B operand from W3t [NP][H] (bf16)
#pragma unroll
  for (int n0 = 16 * wid; n0 < NP; n0 += 16 * WPB) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < H; k0 += 16) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(h2s + r * H + k0 + kq);
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(Pb + g.oW3t + (n0 + r) * H + k0 + kq);
      acc = mfma_bf16(a, b, acc);
    }
    for (int i = 0; i < 4; ++i) {
      const int row = kq + i, col = n0 + r;
      raws[row * NP + col] = acc[i] + P[g.ob3 + col];
    }
  }
  __syncthreads();
}

struct FwdArgs {
  FlowGeom geo;
  int layer;
  int64_t n;
  const float* x_in;       // (n, D)
  float* x_out;            // (n, D)
  float* logdet;           // (n,) accumulated
  float* save_x;           // optional (n, D): copy of the layer input
  // optional bf16 activations, tile-transposed [tile][H][TILE] (sample fastest, zero for the
  // samples past n in the last tile): a backward MFMA operand (four consecutive samples of
  // one column) is then one 8-byte load
  unsigned short* save_h1;
  unsigned short* save_h2;
  float* save_raw;         // optional (n, NPmax)
  int inverse;
};

// NB: the layer's number of transformed dimensions when known at compile time (its NP
// is then a constant and the loops over it unroll), 0 for the run-time geometry.
template <int NB>
__global__ void __launch_bounds__(NT) k_layer(FwdArgs A, const float* __restrict__ P,
                                              const unsigned short* __restrict__ Pb) {
  constexpr int NPC = NB ? ((NB * NPD + 15) / 16) * 16 : 0;
  const LayerGeom& g = A.geo.layer[A.layer];
  const int NP = NPC ? NPC : g.NP, nB = NB ? NB : g.nB;
  const int D = A.geo.D;
  __shared__ float xa[TILE * DMAX];
  __shared__ float xs[TILE * DMAX];
  __shared__ __attribute__((aligned(16))) unsigned short h1s[TILE * H];
  __shared__ __attribute__((aligned(16))) unsigned short h2s[TILE * H];
  extern __shared__ float raws[];
  const int tid = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * TILE;
  for (int q = tid; q < TILE * DMAX; q += NT) {
    const int s = q / DMAX, d = q % DMAX;
    const int64_t e = s0 + s;
    const float v = (e < A.n && d < D) ? A.x_in[e * D + d] : 0.5f;
    xs[q] = v;
    if (A.save_x && e < A.n && d < D) A.save_x[e * D + d] = v;
  }
  __syncthreads();
  for (int q = tid; q < TILE * DMAX; q += NT) {
    const int s = q / DMAX, i = q % DMAX;
    xa[q] = i < g.nA ? xs[s * DMAX + g.A[i]] : 0.f;
  }
  __syncthreads();
  mlp_forward<NPC>(g, P, Pb, xa, h1s, h2s, raws);
  // spline on every (sample, transformed dimension); the log-dets of one sample are
  // summed by one thread afterwards in a fixed order
  __shared__ float lds_ld[TILE * DMAX];
  for (int q = tid; q < TILE * nB; q += NT) {
    const int s = q / nB, b = q % nB;
    const int d = g.B[b];
    const float* raw = raws + s * NP + b * NPD;
    const SplineEval r = A.inverse ? rq_inverse(raw, xs[s * DMAX + d]) : rq_forward(raw, xs[s * DMAX + d]);
    xs[s * DMAX + d] = r.z;
    lds_ld[s * DMAX + b] = r.ld;
  }
  __syncthreads();
  if (tid < TILE) {
    const int64_t e = s0 + tid;
    if (e < A.n) {
      float ld = 0.f;
      for (int b = 0; b < nB; ++b) ld += lds_ld[tid * DMAX + b];
      A.logdet[e] += ld;
    }
  }
  for (int q = tid; q < TILE * D; q += NT) {
    const int s = q / D, d = q % D;
    const int64_t e = s0 + s;
    if (e < A.n) A.x_out[e * D + d] = xs[s * DMAX + d];
  }
  if (A.save_h1) {
    for (int q = tid; q < TILE * H; q += NT) {
      const int j = q / TILE, s = q % TILE;
      const bool ok = s0 + s < A.n;
      A.save_h1[s0 * H + q] = ok ? h1s[s * H + j] : (unsigned short)0;
      A.save_h2[s0 * H + q] = ok ? h2s[s * H + j] : (unsigned short)0;
    }
    for (int q = tid; q < TILE * NP; q += NT) {
      const int64_t e = s0 + q / NP;
      if (e < A.n) A.save_raw[e * A.geo.NPmax + q % NP] = raws[q];
    }
  }
}

// Forward / inverse of one layer with NB (1..NB_LDS) transformed dimensions, persistent:
// each block stages the layer's weights in LDS once (W2^T and W3^T in bf16, rows padded
// to WP = H + 4 elements so that the 16 rows an MFMA operand spans fall in distinct
// banks; W1 and the biases in fp32), then runs the tiles b, b + G, ... of 16 samples
// through the same arithmetic as k_layer.
constexpr int NB_LDS = 6;
constexpr int WP = H + 4;       // padded row pitch (bf16 elements) of the LDS weights and h1/h2

template <int NB>
__global__ void __launch_bounds__(NT) k_layer_lds(FwdArgs A, const float* __restrict__ P,
                                                  const unsigned short* __restrict__ Pb) {
  constexpr int NP = ((NB * NPD + 15) / 16) * 16;
  const LayerGeom& g = A.geo.layer[A.layer];
  const int D = A.geo.D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  __shared__ __attribute__((aligned(16))) unsigned short W2s[H * WP];
  __shared__ __attribute__((aligned(16))) unsigned short W3s[NP * WP];
  __shared__ float W1s[DMAX * H], b1s[H], b2s[H], b3s[NP];
  __shared__ float xa[TILE * DMAX], xs[TILE * DMAX], lds_ld[TILE * DMAX];
  __shared__ __attribute__((aligned(16))) unsigned short h1s[TILE * WP];
  __shared__ __attribute__((aligned(16))) unsigned short h2s[TILE * WP];
  __shared__ float raws[TILE * NP];
  // weights: bf16 rows copied 8 bytes at a time
  for (int q = tid; q < H * (H / 4); q += NT) {
    const int row = q / (H / 4), c = 4 * (q % (H / 4));
    *reinterpret_cast<uint2*>(W2s + row * WP + c) = *reinterpret_cast<const uint2*>(Pb + g.oW2t + row * H + c);
  }
  for (int q = tid; q < NP * (H / 4); q += NT) {
    const int row = q / (H / 4), c = 4 * (q % (H / 4));
    *reinterpret_cast<uint2*>(W3s + row * WP + c) = *reinterpret_cast<const uint2*>(Pb + g.oW3t + row * H + c);
  }
  for (int q = tid; q < g.nA * H; q += NT) W1s[q] = P[g.oW1 + q];
  for (int q = tid; q < H; q += NT) {
    b1s[q] = P[g.ob1 + q];
    b2s[q] = P[g.ob2 + q];
  }
  for (int q = tid; q < NP; q += NT) b3s[q] = P[g.ob3 + q];
  const int64_t ntiles = (A.n + TILE - 1) / TILE;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t s0 = tile * TILE;
    __syncthreads();                 // weights staged / the previous tile is done with the LDS
    for (int q = tid; q < TILE * DMAX; q += NT) {
      const int s = q / DMAX, d = q % DMAX;
      const int64_t e = s0 + s;
      const float v = (e < A.n && d < D) ? A.x_in[e * D + d] : 0.5f;
      xs[q] = v;
      if (A.save_x && e < A.n && d < D) A.save_x[e * D + d] = v;
    }
    __syncthreads();
    for (int q = tid; q < TILE * DMAX; q += NT) {
      const int s = q / DMAX, i = q % DMAX;
      xa[q] = i < g.nA ? xs[s * DMAX + g.A[i]] : 0.f;
    }
    __syncthreads();
    // layer 1 on VALU: 16 x 128 outputs
    for (int q = tid; q < TILE * H; q += NT) {
      const int s = q / H, j = q % H;
      float v = b1s[j];
      for (int i = 0; i < g.nA; ++i) v += xa[s * DMAX + i] * W1s[i * H + j];
      h1s[s * WP + j] = f2bf(fmaxf(v, 0.f));
    }
    __syncthreads();
    // layer 2: h2 = relu(h1 W2 + b2); wave w computes the column tile 16 w
    {
      const int n0 = 16 * wid;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < H; k0 += 16) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(h1s + r * WP + k0 + kq);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(W2s + (n0 + r) * WP + k0 + kq);
        acc = mfma_bf16(a, b, acc);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = kq + i, col = n0 + r;
        h2s[row * WP + col] = f2bf(fmaxf(acc[i] + b2s[col], 0.f));
      }
    }
    __syncthreads();
    // layer 3: raw = h2 W3 + b3
#pragma unroll
    for (int n0 = 16 * wid; n0 < NP; n0 += 16 * WPB) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < H; k0 += 16) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(h2s + r * WP + k0 + kq);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(W3s + (n0 + r) * WP + k0 + kq);
        acc = mfma_bf16(a, b, acc);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = kq + i, col = n0 + r;
        raws[row * NP + col] = acc[i] + b3s[col];
      }
    }
    __syncthreads();
    for (int q = tid; q < TILE * NB; q += NT) {
      const int s = q / NB, b = q % NB;
      const int d = g.B[b];
      const float* raw = raws + s * NP + b * NPD;
      const SplineEval sp = A.inverse ? rq_inverse(raw, xs[s * DMAX + d]) : rq_forward(raw, xs[s * DMAX + d]);
      xs[s * DMAX + d] = sp.z;
      lds_ld[s * DMAX + b] = sp.ld;
    }
    __syncthreads();
    if (tid < TILE) {
      const int64_t e = s0 + tid;
      if (e < A.n) {
        float ld = 0.f;
        for (int b = 0; b < NB; ++b) ld += lds_ld[tid * DMAX + b];
        A.logdet[e] += ld;
      }
    }
    for (int q = tid; q < TILE * D; q += NT) {
      const int s = q / D, d = q % D;
      const int64_t e = s0 + s;
      if (e < A.n) A.x_out[e * D + d] = xs[s * DMAX + d];
    }
    if (A.save_h1) {
      for (int q = tid; q < TILE * H; q += NT) {
        const int j = q / TILE, s = q % TILE;
        const bool ok = s0 + s < A.n;
        A.save_h1[s0 * H + q] = ok ? h1s[s * WP + j] : (unsigned short)0;
        A.save_h2[s0 * H + q] = ok ? h2s[s * WP + j] : (unsigned short)0;
      }
      for (int q = tid; q < TILE * NP; q += NT) {
        const int64_t e = s0 + q / NP;
        if (e < A.n) A.save_raw[e * A.geo.NPmax + q % NP] = raws[q];
      }
    }
  }
}

// ------------------------------------------------------------------ backward
struct BwdArgs {
  FlowGeom geo;
  int layer;
  int64_t n;
  const float* x_in;               // layer input (n, D)
  const unsigned short* h1;        // saved (n, H)
  const unsigned short* h2;        // saved (n, H)
  const float* raw;                // saved (n, NPmax)
  const float* gl;                 // (n,) d objective / d log-det (per sample)
  float* gx;                       // (n, D): in: d/d layer output, out: d/d layer input
  float* partial;                  // (nblocks, layer nparams) weight-gradient partials
};

__global__ void __launch_bounds__(NT) k_layer_bwd(BwdArgs A, const float* __restrict__ P,
                                                  const unsigned short* __restrict__ Pb) {
  const LayerGeom& g = A.geo.layer[A.layer];
  const int D = A.geo.D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = lane >> 4;  // fp32 MFMA: k index = lane/16
  const int64_t s0 = (int64_t)blockIdx.x * TILE;
  __shared__ float xs[TILE * DMAX], gxs[TILE * DMAX];
  __shared__ float h1f[TILE * H], h2f[TILE * H], gh[TILE * H];
  extern __shared__ float go[];   // [TILE][NP]
  for (int q = tid; q < TILE * DMAX; q += NT) {
    const int s = q / DMAX, d = q % DMAX;
    const int64_t e = s0 + s;
    const bool ok = e < A.n && d < D;
    xs[q] = ok ? A.x_in[e * D + d] : 0.5f;
    gxs[q] = ok ? A.gx[e * D + d] : 0.f;
  }
  for (int q = tid; q < TILE * H; q += NT) {
    const int j = q / TILE, s = q % TILE;          // tile-transposed saved activations
    h1f[s * H + j] = bf2f(A.h1[s0 * H + q]);
    h2f[s * H + j] = bf2f(A.h2[s0 * H + q]);
  }
  __syncthreads();
  // spline backward: d raw (into go) and d x_B (replaces the incoming gradient of x_B)
  for (int q = tid; q < TILE * g.NP; q += NT) go[q] = 0.f;
  __syncthreads();
  for (int q = tid; q < TILE * g.nB; q += NT) {
    const int s = q / g.nB, b = q % g.nB;
    const int64_t e = s0 + s;
    if (e >= A.n) continue;
    const int d = g.B[b];
    float graw[NPD];
    const float gxb = rq_backward(A.raw + e * A.geo.NPmax + b * NPD, xs[s * DMAX + d], gxs[s * DMAX + d], A.gl[e],
                                  graw);
    for (int j = 0; j < NPD; ++j) go[s * g.NP + b * NPD + j] = graw[j];
    gxs[s * DMAX + d] = gxb;
  }
  __syncthreads();
  const int nlp = g.ob3 + g.NP - g.oW1;   // parameters of this layer
  float* part = A.partial + (size_t)blockIdx.x * nlp;
  // db3 and dW3 = h2^T go  (M = H, N = NP, K = 16 samples)
  for (int q = tid; q < g.NP; q += NT) {
    float v = 0.f;
    for (int s = 0; s < TILE; ++s) v += go[s * g.NP + q];
    part[g.ob3 - g.oW1 + q] = v;
  }
  const int nt3 = g.NP / 16;
  for (int t = wid; t < (H / 16) * nt3; t += WPB) {
      const int m0 = 16 * (t / nt3), n0 = 16 * (t % nt3);
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int k0 = 0; k0 < TILE; k0 += 4) acc = mfma_f32(h2f[(k0 + kq) * H + m0 + r], go[(k0 + kq) * g.NP + n0 + r], acc);
      for (int i = 0; i < 4; ++i) part[g.oW3 - g.oW1 + (m0 + 4 * kq + i) * g.NP + n0 + r] = acc[i];
    }
  // gh2 = (go W3^T) * relu'(h2)  (M = 16, N = H, K = NP), with the bf16 weights the
  // forward pass used: W3^T[np][j] is the transposed bf16 copy W3t
  for (int n0 = 16 * wid; n0 < H; n0 += 16 * WPB) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.NP; k0 += 4)
      acc = mfma_f32(go[r * g.NP + k0 + kq], bf2f(Pb[g.oW3t + (k0 + kq) * H + n0 + r]), acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * kq + i, col = n0 + r;
      gh[row * H + col] = h2f[row * H + col] > 0.f ? acc[i] : 0.f;
    }
  }
  __syncthreads();
  // db2, dW2 = h1^T gh2
  for (int q = tid; q < H; q += NT) {
    float v = 0.f;
    for (int s = 0; s < TILE; ++s) v += gh[s * H + q];
    part[g.ob2 - g.oW1 + q] = v;
  }
  for (int t = wid; t < (H / 16) * (H / 16); t += WPB) {
      const int m0 = 16 * (t / (H / 16)), n0 = 16 * (t % (H / 16));
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int k0 = 0; k0 < TILE; k0 += 4) acc = mfma_f32(h1f[(k0 + kq) * H + m0 + r], gh[(k0 + kq) * H + n0 + r], acc);
      for (int i = 0; i < 4; ++i) part[g.oW2 - g.oW1 + (m0 + 4 * kq + i) * H + n0 + r] = acc[i];
    }
  // gh1 = (gh2 W2^T) * relu'(h1); W2^T[o][i] is the bf16 copy W2t
  __shared__ float gh1[TILE * H];
  for (int n0 = 16 * wid; n0 < H; n0 += 16 * WPB) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < H; k0 += 4) acc = mfma_f32(gh[r * H + k0 + kq], bf2f(Pb[g.oW2t + (k0 + kq) * H + n0 + r]), acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * kq + i, col = n0 + r;
      gh1[row * H + col] = h1f[row * H + col] > 0.f ? acc[i] : 0.f;
    }
  }
  __syncthreads();
  // db1, dW1 = x_A^T gh1, and d x_A += gh1 W1^T
  for (int q = tid; q < H; q += NT) {
    float v = 0.f;
    for (int s = 0; s < TILE; ++s) v += gh1[s * H + q];
    part[g.ob1 - g.oW1 + q] = v;
  }
  for (int q = tid; q < DMAX * H; q += NT) {
    const int i = q / H, j = q % H;
    float v = 0.f;
    if (i < g.nA)
      for (int s = 0; s < TILE; ++s) v += xs[s * DMAX + g.A[i]] * gh1[s * H + j];
    part[q] = v;   // W1 block is the first DMAX*H entries of the layer
  }
  for (int q = tid; q < TILE * g.nA; q += NT) {
    const int s = q / g.nA, i = q % g.nA;
    float v = 0.f;
    for (int j = 0; j < H; ++j) v += gh1[s * H + j] * P[g.oW1 + i * H + j];
    gxs[s * DMAX + g.A[i]] += v;
  }
  __syncthreads();
  for (int q = tid; q < TILE * D; q += NT) {
    const int s = q / D, d = q % D;
    const int64_t e = s0 + s;
    if (e < A.n) A.gx[e * D + d] = gxs[s * DMAX + d];
  }
}

// Backward of one layer in two kernels, for layers whose transformed set has NB (1..8)
// dimensions (NP = pad16(NB (3K+1)) is then a compile-time constant, so every loop over
// it unrolls).  Both run their GEMMs on the bf16 matrix cores with fp32 accumulation
// (v_mfma_f32_16x16x16bf16): the activations and weights are the bf16 values of the
// forward pass, the deltas are rounded to bf16 like them.
//  * k_layer_bwd_data: per 16-sample tile (one per block), the spline backward and the
//    two input-gradient products gh2 = (go W3^T) relu'(h2), gh1 = (gh2 W2^T) relu'(h1),
//    the input gradient gx, and the deltas go, gh2, gh1 written tile-transposed
//    ([tile][column][sample], like the saved activations);
//  * k_layer_wgrad: the weight gradients as GEMMs with the samples as K, over splits of
//    WG_TILES tiles: block (split, y < 8) computes rows 16 y .. 16 y + 15 of dW2 = h1^T gh2
//    and dW3 = h2^T go (wave w: column tiles w, w + 8, ...); block (split, 8) the bias sums
//    as products with a row of ones and dW1 = x_A^T gh1 (fp32, from the exact inputs).
//    One partial per split, reduced in a fixed order.
constexpr int WG_TILES = 4;                    // tiles (of 16 samples) per weight-gradient split
constexpr int WG_SPLIT = WG_TILES * TILE;      // samples per split

__device__ __forceinline__ bf16x4 ld_bf16x4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}

template <int NB>
__global__ void __launch_bounds__(NT) k_layer_bwd_data(BwdArgs A, const float* __restrict__ P,
                                                       const unsigned short* __restrict__ Pb,
                                                       unsigned short* __restrict__ d_o,
                                                       unsigned short* __restrict__ d_h2,
                                                       unsigned short* __restrict__ d_h1) {
  constexpr int NP = ((NB * NPD + 15) / 16) * 16;
  constexpr int NPP = NP + 4;                  // padded bf16 row of go (distinct banks per row)
  static_assert(H / 16 == WPB, "one column tile of gh2 / gh1 per wave");
  const LayerGeom& g = A.geo.layer[A.layer];
  const int D = A.geo.D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;   // bf16 MFMA: four consecutive k per lane
  const int n0 = 16 * wid;                         // this wave's column tile of gh2 / gh1
  const int64_t tile = blockIdx.x, s0 = tile * TILE;
  __shared__ float xs[TILE * DMAX], gxs[TILE * DMAX];
  __shared__ __attribute__((aligned(16))) unsigned short h1b[TILE * WP], h2b[TILE * WP], ghb[TILE * WP];
  __shared__ __attribute__((aligned(16))) unsigned short gob[TILE * NPP];
  __shared__ float go[TILE * NP], gh1[TILE * H], W1s[DMAX * H];
  for (int q = tid; q < TILE * DMAX; q += NT) {
    const int s = q / DMAX, d = q % DMAX;
    const int64_t e = s0 + s;
    const bool ok = e < A.n && d < D;
    xs[q] = ok ? A.x_in[e * D + d] : 0.5f;
    gxs[q] = ok ? A.gx[e * D + d] : 0.f;
  }
  for (int q = tid; q < g.nA * H; q += NT) W1s[q] = P[g.oW1 + q];
  for (int q = tid; q < TILE * H; q += NT) {
    const int j = q / TILE, s = q % TILE;        // saved activations are tile-transposed
    h1b[s * WP + j] = A.h1[s0 * H + q];
    h2b[s * WP + j] = A.h2[s0 * H + q];
  }
  for (int q = tid; q < TILE * NP; q += NT) go[q] = 0.f;
  __syncthreads();
  // spline backward: d raw (into go) and d x_B (replaces the incoming gradient of x_B),
  // written straight into LDS (a private array indexed by the bin would live in scratch)
  for (int q = tid; q < TILE * NB; q += NT) {
    const int s = q / NB, b = q % NB;
    const int64_t e = s0 + s;
    if (e >= A.n) continue;
    const int d = g.B[b];
    gxs[s * DMAX + d] = rq_backward(A.raw + e * A.geo.NPmax + b * NPD, xs[s * DMAX + d], gxs[s * DMAX + d],
                                    A.gl[e], go + s * NP + b * NPD);
  }
  __syncthreads();
  // go in bf16: the A operand of gh2 (LDS) and the delta of dW3 / b3 (global, tile-transposed)
  for (int q = tid; q < TILE * NP; q += NT) {
    const int c = q / TILE, s = q % TILE;
    const unsigned short v = f2bf(go[s * NP + c]);
    gob[s * NPP + c] = v;
    d_o[s0 * NP + q] = v;
  }
  __syncthreads();
  // gh2 = (go W3^T) * relu'(h2): B[k = np][n = h] = W3[h][np], the row h of W3n
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < NP; k0 += 16)
      acc = mfma_bf16(ld_bf16x4(gob + r * NPP + k0 + kq), ld_bf16x4(Pb + g.oW3n + (n0 + r) * NP + k0 + kq), acc);
    unsigned short out[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = kq + i, col = n0 + r;
      const unsigned short v = f2bf(bf2f(h2b[row * WP + col]) > 0.f ? acc[i] : 0.f);
      ghb[row * WP + col] = v;
      out[i] = v;
    }
    // the four samples of column n0 + r: one 8-byte store
    *reinterpret_cast<uint2*>(d_h2 + s0 * H + (n0 + r) * TILE + kq) =
        make_uint2((unsigned)out[0] | ((unsigned)out[1] << 16), (unsigned)out[2] | ((unsigned)out[3] << 16));
  }
  __syncthreads();
  // gh1 = (gh2 W2^T) * relu'(h1): B[k = o][n = i] = W2[i][o], the row i of W2n
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < H; k0 += 16)
      acc = mfma_bf16(ld_bf16x4(ghb + r * WP + k0 + kq), ld_bf16x4(Pb + g.oW2n + (n0 + r) * H + k0 + kq), acc);
    unsigned short out[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = kq + i, col = n0 + r;
      const float v = bf2f(h1b[row * WP + col]) > 0.f ? acc[i] : 0.f;
      gh1[row * H + col] = v;
      out[i] = f2bf(v);
    }
    *reinterpret_cast<uint2*>(d_h1 + s0 * H + (n0 + r) * TILE + kq) =
        make_uint2((unsigned)out[0] | ((unsigned)out[1] << 16), (unsigned)out[2] | ((unsigned)out[3] << 16));
  }
  __syncthreads();
  // d x_A += gh1 W1^T (fp32): 8 lanes per (sample, input), 16 terms each, summed by shuffles
  // within aligned groups of 8 lanes (the summation order is fixed)
  for (int q0 = 0; q0 < TILE * DMAX * 8; q0 += NT) {
    const int q = q0 + tid, o = q >> 3, part = q & 7;
    const int s = o / DMAX, i = o % DMAX;
    float v = 0.f;
    if (i < g.nA) {
#pragma unroll
      for (int j = 0; j < H / 8; ++j) v += gh1[s * H + part * (H / 8) + j] * W1s[i * H + part * (H / 8) + j];
    }
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    if (part == 0 && i < g.nA) gxs[s * DMAX + g.A[i]] += v;
  }
  __syncthreads();
  for (int q = tid; q < TILE * D; q += NT) {
    const int s = q / D, d = q % D;
    const int64_t e = s0 + s;
    if (e < A.n) A.gx[e * D + d] = gxs[s * DMAX + d];
  }
}

template <int NB>
__global__ void __launch_bounds__(NT) k_layer_wgrad(BwdArgs A, const unsigned short* __restrict__ d_o,
                                                    const unsigned short* __restrict__ d_h2,
                                                    const unsigned short* __restrict__ d_h1) {
  constexpr int NP = ((NB * NPD + 15) / 16) * 16;
  constexpr int NT3 = NP / 16;
  constexpr int J3 = (NT3 + WPB - 1) / WPB;
  static_assert(DMAX == 16, "dW1 is one 16-row tile");
  const LayerGeom& g = A.geo.layer[A.layer];
  const int D = A.geo.D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int64_t ntiles = (A.n + TILE - 1) / TILE;
  const int64_t t0 = (int64_t)blockIdx.x * WG_TILES;
  const int nt = (int)(t0 + WG_TILES < ntiles ? WG_TILES : ntiles - t0);   // tiles of this split
  const int nlp = g.ob3 + NP - g.oW1;
  float* part = A.partial + (size_t)blockIdx.x * nlp;
  // a wave whose second dW3 column tile does not exist repeats its last one, unstored
  int jc[J3];
#pragma unroll
  for (int u = 0; u < J3; ++u) jc[u] = wid + WPB * u < NT3 ? wid + WPB * u : NT3 - 1;
  f32x4 acc3[J3], acc2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < J3; ++u) acc3[u] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (blockIdx.y < H / 16) {
    const int m0 = 16 * blockIdx.y;
    // every operand of the split is loaded before the MFMAs (tiles past the end repeat the
    // last one with a zero A operand)
    bf16x4 a3[WG_TILES], a2[WG_TILES], b2[WG_TILES], b3[WG_TILES][J3];
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t) {
      const int64_t tt = t0 + (t < nt ? t : nt - 1);
      a3[t] = ld_bf16x4(A.h2 + tt * TILE * H + (m0 + r) * TILE + kq);
      a2[t] = ld_bf16x4(A.h1 + tt * TILE * H + (m0 + r) * TILE + kq);
      b2[t] = ld_bf16x4(d_h2 + tt * TILE * H + (16 * wid + r) * TILE + kq);
#pragma unroll
      for (int u = 0; u < J3; ++u) b3[t][u] = ld_bf16x4(d_o + tt * TILE * NP + (16 * jc[u] + r) * TILE + kq);
    }
    const bf16x4 zero = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t) {
      const bf16x4 x3 = t < nt ? a3[t] : zero, x2 = t < nt ? a2[t] : zero;
#pragma unroll
      for (int u = 0; u < J3; ++u) acc3[u] = mfma_bf16(x3, b3[t][u], acc3[u]);
      acc2 = mfma_bf16(x2, b2[t], acc2);
    }
#pragma unroll
    for (int u = 0; u < J3; ++u) {
      const int j = wid + WPB * u;
      if (j < NT3)
#pragma unroll
        for (int i = 0; i < 4; ++i) part[g.oW3 - g.oW1 + (m0 + kq + i) * NP + 16 * j + r] = acc3[u][i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) part[g.oW2 - g.oW1 + (m0 + kq + i) * H + 16 * wid + r] = acc2[i];
  } else {
    // bias sums: ones^T B, every row of which is the column sum of B (row 0 is stored)
    const short one = (short)0x3f80;            // 1.0 in bf16
    f32x4 acc1 = {0.f, 0.f, 0.f, 0.f}, accw1 = {0.f, 0.f, 0.f, 0.f};
    bf16x4 bh1[WG_TILES], bh2[WG_TILES], bo[WG_TILES][J3];
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t) {
      const int64_t tt = t0 + (t < nt ? t : nt - 1);
      bh1[t] = ld_bf16x4(d_h1 + tt * TILE * H + (16 * wid + r) * TILE + kq);
      bh2[t] = ld_bf16x4(d_h2 + tt * TILE * H + (16 * wid + r) * TILE + kq);
#pragma unroll
      for (int u = 0; u < J3; ++u) bo[t][u] = ld_bf16x4(d_o + tt * TILE * NP + (16 * jc[u] + r) * TILE + kq);
    }
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t) {
      const bf16x4 ones = t < nt ? bf16x4{one, one, one, one} : bf16x4{0, 0, 0, 0};
      acc1 = mfma_bf16(ones, bh1[t], acc1);
      acc2 = mfma_bf16(ones, bh2[t], acc2);
#pragma unroll
      for (int u = 0; u < J3; ++u) acc3[u] = mfma_bf16(ones, bo[t][u], acc3[u]);
    }
    // dW1 = x_A^T gh1 on the fp32 matrix cores (A[r][k] = x of sample k, input r < nA)
    const int kf = lane >> 4;                    // fp32 MFMA: one k per lane quarter
    const bool xrow = r < g.nA;
    const int da = xrow ? g.A[r] : 0;
    float xv[WG_TILES][4], gv[WG_TILES][4];
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t) {
      const int64_t tt = t0 + (t < nt ? t : nt - 1);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t e = tt * TILE + 4 * k + kf;
        xv[t][k] = A.x_in[(e < A.n ? e : A.n - 1) * D + da];
        gv[t][k] = bf2f(d_h1[tt * TILE * H + (16 * wid + r) * TILE + 4 * k + kf]);
      }
    }
#pragma unroll
    for (int t = 0; t < WG_TILES; ++t)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t e = (t0 + t) * TILE + 4 * k + kf;
        const bool ok = t < nt && e < A.n && xrow;
        accw1 = mfma_f32(ok ? xv[t][k] : 0.f, gv[t][k], accw1);
      }
    // fp32 MFMA output: D[4 kf + i][r] in element i
#pragma unroll
    for (int i = 0; i < 4; ++i) part[(4 * kf + i) * H + 16 * wid + r] = accw1[i];
    if (kq == 0) {
      part[g.ob1 - g.oW1 + 16 * wid + r] = acc1[0];
      part[g.ob2 - g.oW1 + 16 * wid + r] = acc2[0];
#pragma unroll
      for (int u = 0; u < J3; ++u) {
        const int j = wid + WPB * u;
        if (j < NT3) part[g.ob3 - g.oW1 + 16 * j + r] = acc3[u][0];
      }
    }
  }
}

// fixed-order reduction of the per-block partials into the gradient vector
// Sum of the per-block partial gradients: 64 parameters per block, the partials of each
// split over 4 waves (Kahan sums over blocks w, w+4, ...), combined in a fixed order in
// LDS, so the result does not depend on scheduling.
__global__ void __launch_bounds__(256) k_reduce_grad(const float* __restrict__ partial, int nblocks, int nlp,
                                                     float* __restrict__ grad) {
  __shared__ float ls[4][64], lc[4][64];
  const int pi = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + pi;
  float s = 0.f, c = 0.f;
  if (i < nlp) {
    // the loads of 8 blocks are issued before their compensated adds, which keep the
    // order b = sl, sl + 4, ...
    for (int b0 = sl; b0 < nblocks; b0 += 32) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int b = b0 + 4 * u;
        v[u] = b < nblocks ? partial[(size_t)b * nlp + i] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (b0 + 4 * u < nblocks) {
          const float y = v[u] - c;
          const float t = s + y;
          c = (t - s) - y;
          s = t;
        }
      }
    }
  }
  ls[sl][pi] = s;
  lc[sl][pi] = c;
  __syncthreads();
  if (sl == 0 && i < nlp) {
    float tot = 0.f, cc = 0.f;
    for (int w = 0; w < 4; ++w) {
      const float y = (ls[w][pi] - lc[w][pi]) - cc;
      const float t = tot + y;
      cc = (t - tot) - y;
      tot = t;
    }
    grad[i] = tot;
  }
}

// Minibatch draw with probability proportional to the training weights: sample i takes
// the first index k with cdf[k] >= r[i] (cdf = normalised running sum of the weights),
// and copies that point's coordinates and integrand value.
__global__ void k_minibatch(const double* __restrict__ cdf, int64_t n, const double* __restrict__ r, int64_t batch,
                            const float* __restrict__ y, int D, const double* __restrict__ f, float* __restrict__ yb,
                            double* __restrict__ fb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  const double u = r[i];
  int64_t lo = 0, hi = n - 1;          // invariant: the answer lies in [lo, hi]
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  for (int d = 0; d < D; ++d) yb[i * D + d] = y[lo * D + d];
  if (f) fb[i] = f[lo];
}

// Per-sample seed of the backward pass for the self-normalised objective
// -sum_i v_i log q_i / sum_i v_i with v_i = w_i (f_i / q_i)^beta (w_i = 1 if w is null):
// gl_i = -v_i / sum v.  One block; the maximum (for a safe exp) and the sum are reduced
// in LDS in a fixed order.
constexpr int SEED_T = 1024;
__global__ void __launch_bounds__(SEED_T) k_seed(const double* __restrict__ w, const double* __restrict__ f,
                                                 const float* __restrict__ ld, int n, float beta, float* __restrict__ gl) {
  __shared__ double red[SEED_T];
  const int t = threadIdx.x;
  // f_i <= 0 (or not a number) gives v_i = 0
  auto lv = [&](int i) { return f[i] > 0.0 ? (double)beta * (log(f[i]) - (double)ld[i]) : -HUGE_VAL; };
  double mx = -HUGE_VAL;
  if (beta != 0.f)
    for (int i = t; i < n; i += SEED_T) mx = fmax(mx, lv(i));
  red[t] = mx;
  __syncthreads();
  for (int s = SEED_T / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = fmax(red[t], red[t + s]);
    __syncthreads();
  }
  mx = beta != 0.f ? red[0] : 0.0;
  if (!(mx > -HUGE_VAL)) mx = 0.0;     // every f_i = 0: all v_i = 0, the seed is zero
  __syncthreads();
  double sum = 0.0;
  for (int i = t; i < n; i += SEED_T) sum += (w ? w[i] : 1.0) * (beta != 0.f ? exp(lv(i) - mx) : 1.0);
  red[t] = sum;
  __syncthreads();
  for (int s = SEED_T / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double tot = red[0];
  for (int i = t; i < n; i += SEED_T) {
    const double v = (w ? w[i] : 1.0) * (beta != 0.f ? exp(lv(i) - mx) : 1.0);
    gl[i] = (float)(tot > 0.0 ? -v / tot : 0.0);
  }
}

// Adam step on the fp32 master parameters, then refresh the transposed bf16 copies.
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float bc1, float bc2,
                       float clip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = grad[i];
  if (!isfinite(gi)) gi = 0.f;
  gi = fminf(fmaxf(gi, -clip), clip);
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  P[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

__global__ void k_cast(FlowGeom geo, const float* __restrict__ P, unsigned short* __restrict__ Pb) {
  const int l = blockIdx.y;
  const LayerGeom& g = geo.layer[l];
  const int tot = H * H + g.NP * H;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += gridDim.x * blockDim.x) {
    if (q < H * H) {
      const int o = q / H, i = q % H;       // W2t[o][i] = W2[i][o]
      Pb[g.oW2t + q] = f2bf(P[g.oW2 + i * H + o]);
      Pb[g.oW2n + q] = f2bf(P[g.oW2 + q]);  // W2n = W2
    } else {
      const int qq = q - H * H;
      const int o = qq / H, i = qq % H;     // W3t[o][i] = W3[i][o]
      Pb[g.oW3t + qq] = f2bf(P[g.oW3 + i * g.NP + o]);
      Pb[g.oW3n + qq] = f2bf(P[g.oW3 + qq]);  // W3n = W3
    }
  }
}

}  // namespace flow
}  // namespace mfx

using namespace mfx;
using namespace mfx::flow;

// Layers with 1..NB_TEMPLATED transformed dimensions run the kernels compiled for their
// geometry (NP a constant); wider layers the run-time-geometry kernels.
constexpr int NB_TEMPLATED = 8;

static void launch_layer(int nb, dim3 grid, size_t shmem, hipStream_t s, const FwdArgs& A, const float* P,
                         const unsigned short* Pb) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  if (nb >= 1 && nb <= NB_LDS && grid.x <= 8u * (unsigned)cus) {
    // training-size batches: persistent blocks with the weights in LDS, one per CU at
    // most, each running a few tiles (the weights are read once per block).  Large
    // batches (sampling) keep one block per tile, which holds more waves per SIMD.
    const dim3 gp(grid.x < (unsigned)cus ? grid.x : (unsigned)cus);
    switch (nb) {
      case 1: hipLaunchKernelGGL(k_layer_lds<1>, gp, dim3(NT), 0, s, A, P, Pb); break;
      case 2: hipLaunchKernelGGL(k_layer_lds<2>, gp, dim3(NT), 0, s, A, P, Pb); break;
      case 3: hipLaunchKernelGGL(k_layer_lds<3>, gp, dim3(NT), 0, s, A, P, Pb); break;
      case 4: hipLaunchKernelGGL(k_layer_lds<4>, gp, dim3(NT), 0, s, A, P, Pb); break;
      case 5: hipLaunchKernelGGL(k_layer_lds<5>, gp, dim3(NT), 0, s, A, P, Pb); break;
      default: hipLaunchKernelGGL(k_layer_lds<6>, gp, dim3(NT), 0, s, A, P, Pb); break;
    }
    return;
  }
  switch (nb) {
    case 1: hipLaunchKernelGGL(k_layer<1>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 2: hipLaunchKernelGGL(k_layer<2>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 3: hipLaunchKernelGGL(k_layer<3>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 4: hipLaunchKernelGGL(k_layer<4>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 5: hipLaunchKernelGGL(k_layer<5>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 6: hipLaunchKernelGGL(k_layer<6>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 7: hipLaunchKernelGGL(k_layer<7>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    case 8: hipLaunchKernelGGL(k_layer<8>, grid, dim3(NT), shmem, s, A, P, Pb); break;
    default: hipLaunchKernelGGL(k_layer<0>, grid, dim3(NT), shmem, s, A, P, Pb); break;
  }
}

// The backward of a layer with 1..NB_TEMPLATED transformed dimensions: the data-gradient
// kernel (one block per 16-sample tile), then the weight-gradient GEMMs (one block row per
// WG_SPLIT samples).  The deltas live at the start of the partial workspace, the
// per-split partials after them.
// The backward of a layer with 1..NB_TEMPLATED transformed dimensions: the data-gradient
// kernel (one block per 16-sample tile), then the weight-gradient GEMMs (one block row per
// WG_TILES tiles).  The deltas live at the start of the partial workspace, the per-split
// partials after them.
static void launch_layer_bwd_split(int nb, int64_t n, hipStream_t s, const BwdArgs& A, const float* P,
                                   const unsigned short* Pb, unsigned short* d_o, unsigned short* d_h2,
                                   unsigned short* d_h1) {
  const dim3 gd((unsigned)((n + TILE - 1) / TILE)), gw((unsigned)((n + WG_SPLIT - 1) / WG_SPLIT), H / 16 + 1);
#define MFX_BWD_CASE(K)                                                                                    \
  case K:                                                                                                 \
    hipLaunchKernelGGL(k_layer_bwd_data<K>, gd, dim3(NT), 0, s, A, P, Pb, d_o, d_h2, d_h1);               \
    hipLaunchKernelGGL(k_layer_wgrad<K>, gw, dim3(NT), 0, s, A, (const unsigned short*)d_o,                \
                       (const unsigned short*)d_h2, (const unsigned short*)d_h1);                         \
    break;
  switch (nb) {
    MFX_BWD_CASE(1)
    MFX_BWD_CASE(2)
    MFX_BWD_CASE(3)
    MFX_BWD_CASE(4)
    MFX_BWD_CASE(5)
    MFX_BWD_CASE(6)
    MFX_BWD_CASE(7)
    default: MFX_BWD_CASE(8)
  }
#undef MFX_BWD_CASE
}

static int64_t ntiles_of(int64_t n) { return (n + TILE - 1) / TILE; }
static size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// Saved state of one layer (training), each piece 256-byte aligned: x (n, D) and raw
// (n, NPmax) in f32, then the bf16 activations h1 and h2 tile-transposed over whole tiles.
struct SaveLayout {
  size_t raw, h1, h2, bytes;   // byte offsets from the layer's base, and the layer's size
};
static SaveLayout save_layout(const FlowGeom& G, int64_t n) {
  SaveLayout s;
  s.raw = align256((size_t)n * G.D * 4);
  s.h1 = s.raw + align256((size_t)n * G.NPmax * 4);
  s.h2 = s.h1 + (size_t)ntiles_of(n) * TILE * H * 2;
  s.bytes = s.h2 + (size_t)ntiles_of(n) * TILE * H * 2;
  return s;
}
static size_t save_per_layer(const FlowGeom& G, int64_t n) { return save_layout(G, n).bytes; }

static int make_geom(int D, int L, const int* masks, FlowGeom& G) {
  if (D < 2 || D > DMAX || L < 1 || L > 16) return fail(-2, "flow needs 2 <= D <= 16 and 1 <= L <= 16%s", "");
  G.D = D;
  G.L = L;
  int np = 0, nb = 0, npmax = 0;
  for (int l = 0; l < L; ++l) {
    LayerGeom& g = G.layer[l];
    g.nA = g.nB = 0;
    for (int d = 0; d < D; ++d) {
      if (masks[l * D + d]) g.B[g.nB++] = d;
      else g.A[g.nA++] = d;
    }
    if (g.nA < 1 || g.nB < 1) return fail(-2, "every flow layer needs conditioning and transformed dims%s", "");
    g.NP = ((g.nB * NPD + 15) / 16) * 16;
    npmax = g.NP > npmax ? g.NP : npmax;
    g.oW1 = np;
    g.ob1 = g.oW1 + DMAX * H;
    g.oW2 = g.ob1 + H;
    g.ob2 = g.oW2 + H * H;
    g.oW3 = g.ob2 + H;
    g.ob3 = g.oW3 + H * g.NP;
    np = g.ob3 + g.NP;
    g.oW2t = nb;
    g.oW3t = nb + H * H;
    g.oW2n = g.oW3t + g.NP * H;
    g.oW3n = g.oW2n + H * H;
    nb += 2 * (H * H + g.NP * H);
  }
  G.nparams = np;
  G.nbf16 = nb;
  G.NPmax = npmax;
  return 0;
}

extern "C" {

int mfx_flow_sizes(int D, int L, const int* masks, int64_t* sizes) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  sizes[0] = G.nparams;
  sizes[1] = G.nbf16;
  sizes[2] = G.NPmax;
  // per-layer parameter offset and count
  for (int l = 0; l < L; ++l) {
    sizes[3 + 2 * l] = G.layer[l].oW1;
    sizes[4 + 2 * l] = G.layer[l].ob3 + G.layer[l].NP - G.layer[l].oW1;
  }
  return 0;
}

int mfx_flow_cast(int D, int L, const int* masks, const float* params, uint16_t* params_bf16, void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cast, dim3(64, L), dim3(256), 0, (hipStream_t)stream, G, params, params_bf16);
  return hip_status(hipGetLastError());
}

int mfx_flow_apply(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16, int64_t n,
                   const float* x_in, float* x_out, float* logdet, int inverse, float* save, void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  if (n < 0) return fail(-2, "n must be >= 0%s", "");
  if (n == 0) return 0;
  if (!params || !params_bf16 || !x_in || !x_out || !logdet) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((n + TILE - 1) / TILE);
  // per layer saved state (training): x_in (n, D), raw (n, NPmax) f32, then h1, h2 bf16
  // tile-transposed over whole tiles
  const size_t per_layer = save_per_layer(G, n);
  const float* cur = x_in;
  for (int step = 0; step < L; ++step) {
    const int l = inverse ? L - 1 - step : step;
    FwdArgs A;
    A.geo = G;
    A.layer = l;
    A.n = n;
    A.x_in = cur;
    A.x_out = x_out;
    A.logdet = logdet;
    A.inverse = inverse;
    A.save_x = nullptr;
    A.save_h1 = A.save_h2 = nullptr;
    A.save_raw = nullptr;
    if (save && !inverse) {
      char* base = (char*)save + per_layer * l;
      const SaveLayout sl = save_layout(G, n);
      A.save_x = (float*)base;   // the kernel keeps the layer input for the backward pass
      A.save_raw = (float*)(base + sl.raw);
      A.save_h1 = (unsigned short*)(base + sl.h1);
      A.save_h2 = (unsigned short*)(base + sl.h2);
    }
    const size_t shmem = (size_t)TILE * G.layer[l].NP * 4;
    launch_layer(G.layer[l].nB, dim3(nblk), shmem, s, A, params, (const unsigned short*)params_bf16);
    cur = x_out;
  }
  return hip_status(hipGetLastError());
}

int64_t mfx_flow_save_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G)) return -1;
  return (int64_t)L * (int64_t)save_per_layer(G, n);
}

int64_t mfx_flow_partial_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G)) return -1;
  int maxlp = 0;
  for (int l = 0; l < L; ++l) {
    const int lp = G.layer[l].ob3 + G.layer[l].NP - G.layer[l].oW1;
    maxlp = lp > maxlp ? lp : maxlp;
  }
  // the larger of the two layouts: per-tile partials (run-time-geometry backward), or the
  // deltas of the batch (NP + 2 H floats per sample) followed by one partial per
  // WG_SPLIT samples (the split backward)
  int64_t bytes = ((n + TILE - 1) / TILE) * (int64_t)maxlp * 4;
  for (int l = 0; l < L; ++l) {
    const LayerGeom& g = G.layer[l];
    if (g.nB < 1 || g.nB > NB_TEMPLATED) continue;
    const int64_t lp = g.ob3 + g.NP - g.oW1;
    const int64_t b = ((n + TILE - 1) / TILE) * TILE * (g.NP + 2 * H) * 2 + (n + WG_SPLIT - 1) / WG_SPLIT * lp * 4;
    bytes = b > bytes ? b : bytes;
  }
  return bytes;
}

int mfx_flow_backward(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16, int64_t n,
                      const float* save, const float* gl, float* gx, float* partial, float* grad, void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  if (n <= 0) return n < 0 ? fail(-2, "n must be >= 0%s", "") : 0;
  if (!params || !params_bf16 || !save || !gl || !gx || !partial || !grad) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((n + TILE - 1) / TILE);
  const size_t per_layer = save_per_layer(G, n);
  for (int l = L - 1; l >= 0; --l) {
    const char* base = (const char*)save + per_layer * l;
    const SaveLayout sl = save_layout(G, n);
    BwdArgs A;
    A.geo = G;
    A.layer = l;
    A.n = n;
    A.x_in = (const float*)base;
    A.raw = (const float*)(base + sl.raw);
    A.h1 = (const unsigned short*)(base + sl.h1);
    A.h2 = (const unsigned short*)(base + sl.h2);
    A.gl = gl;
    A.gx = gx;
    A.partial = partial;
    const int nlp = G.layer[l].ob3 + G.layer[l].NP - G.layer[l].oW1;
    int nparts = nblk;
    const int nb = G.layer[l].nB;
    if (nb >= 1 && nb <= NB_TEMPLATED) {
      // bf16 deltas over whole tiles (NP + 2 H per sample) first, then one partial per split
      const size_t nt16 = (size_t)((n + TILE - 1) / TILE) * TILE;
      unsigned short* d_o = (unsigned short*)partial;
      unsigned short* d_h2 = d_o + nt16 * G.layer[l].NP;
      unsigned short* d_h1 = d_h2 + nt16 * H;
      nparts = (int)((n + WG_SPLIT - 1) / WG_SPLIT);
      A.partial = (float*)(d_h1 + nt16 * H);
      launch_layer_bwd_split(nb, n, s, A, params, (const unsigned short*)params_bf16, d_o, d_h2, d_h1);
      hipLaunchKernelGGL(k_reduce_grad, dim3((nlp + 63) / 64), dim3(256), 0, s, A.partial, nparts, nlp,
                         grad + G.layer[l].oW1);
      continue;
    } else {
      const size_t shmem = (size_t)TILE * G.layer[l].NP * 4;
      hipLaunchKernelGGL(k_layer_bwd, dim3(nblk), dim3(NT), shmem, s, A, params, (const unsigned short*)params_bf16);
    }
    hipLaunchKernelGGL(k_reduce_grad, dim3((nlp + 63) / 64), dim3(256), 0, s, partial, nparts, nlp,
                       grad + G.layer[l].oW1);
  }
  return hip_status(hipGetLastError());
}

int mfx_flow_minibatch(const double* cdf, int64_t n, const double* r, int64_t batch, const float* y, int D,
                       const double* f, float* yb, double* fb, void* stream) {
  if (n <= 0 || batch < 0 || D < 1 || D > DMAX) return fail(-2, "n > 0, batch >= 0 and 1 <= D <= 16 required%s", "");
  if (!cdf || !r || !y || !yb || (f && !fb)) return fail(-3, "null pointer%s", "");
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_minibatch, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, (hipStream_t)stream, cdf, n, r,
                     batch, y, D, f, yb, fb);
  return hip_status(hipGetLastError());
}

int mfx_flow_seed(const double* w, const double* f, const float* logq, int64_t n, float beta, float* gl, void* stream) {
  if (n <= 0 || n > (1 << 24)) return fail(-2, "0 < n <= 2^24 required%s", "");
  if (!logq || !gl || (beta != 0.f && !f)) return fail(-3, "null pointer (f is needed for beta != 0)%s", "");
  hipLaunchKernelGGL(k_seed, dim3(1), dim3(SEED_T), 0, (hipStream_t)stream, w, f, logq, (int)n, beta, gl);
  return hip_status(hipGetLastError());
}

int mfx_flow_adam(float* params, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1,
                  float beta2, float eps, int step, float clip, void* stream) {
  if (n <= 0 || step < 1) return fail(-2, "n > 0 and step >= 1 required%s", "");
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, grad, m1,
                     m2, (int)n, lr, beta1, beta2, eps, bc1, bc2, clip);
  return hip_status(hipGetLastError());
}

}  // extern "C"
