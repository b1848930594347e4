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
  // offsets (in bf16 elements) of the transposed bf16 copies W2t [H][H], W3t [NP][H]
  int oW2t, oW3t;
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
__device__ void mlp_forward(const LayerGeom& g, const float* __restrict__ P, const unsigned short* __restrict__ Pb,
                            const float* xa, unsigned short* h1s, unsigned short* h2s, float* raws) {
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
  for (int n0 = 16 * wid; n0 < g.NP; n0 += 16 * WPB) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < H; k0 += 16) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(h2s + r * H + k0 + kq);
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(Pb + g.oW3t + (n0 + r) * H + k0 + kq);
      acc = mfma_bf16(a, b, acc);
    }
    for (int i = 0; i < 4; ++i) {
      const int row = kq + i, col = n0 + r;
      raws[row * g.NP + col] = acc[i] + P[g.ob3 + col];
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
  unsigned short* save_h1; // optional (n, H) bf16
  unsigned short* save_h2; // optional (n, H) bf16
  float* save_raw;         // optional (n, NPmax)
  int inverse;
};

__global__ void __launch_bounds__(NT) k_layer(FwdArgs A, const float* __restrict__ P,
                                              const unsigned short* __restrict__ Pb) {
  const LayerGeom& g = A.geo.layer[A.layer];
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
  mlp_forward(g, P, Pb, xa, h1s, h2s, raws);
  // spline on every (sample, transformed dimension); the log-dets of one sample are
  // summed by one thread afterwards in a fixed order
  __shared__ float lds_ld[TILE * DMAX];
  for (int q = tid; q < TILE * g.nB; q += NT) {
    const int s = q / g.nB, b = q % g.nB;
    const int d = g.B[b];
    const float* raw = raws + s * g.NP + b * NPD;
    const SplineEval r = A.inverse ? rq_inverse(raw, xs[s * DMAX + d]) : rq_forward(raw, xs[s * DMAX + d]);
    xs[s * DMAX + d] = r.z;
    lds_ld[s * DMAX + b] = r.ld;
  }
  __syncthreads();
  if (tid < TILE) {
    const int64_t e = s0 + tid;
    if (e < A.n) {
      float ld = 0.f;
      for (int b = 0; b < g.nB; ++b) ld += lds_ld[tid * DMAX + b];
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
      const int64_t e = s0 + q / H;
      if (e < A.n) {
        A.save_h1[e * H + q % H] = h1s[q];
        A.save_h2[e * H + q % H] = h2s[q];
      }
    }
    for (int q = tid; q < TILE * g.NP; q += NT) {
      const int64_t e = s0 + q / g.NP;
      if (e < A.n) A.save_raw[e * A.geo.NPmax + q % g.NP] = raws[q];
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
    const int64_t e = s0 + q / H;
    h1f[q] = e < A.n ? bf2f(A.h1[e * H + q % H]) : 0.f;
    h2f[q] = e < A.n ? bf2f(A.h2[e * H + q % H]) : 0.f;
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
    } else {
      const int qq = q - H * H;
      const int o = qq / H, i = qq % H;     // W3t[o][i] = W3[i][o]
      Pb[g.oW3t + qq] = f2bf(P[g.oW3 + i * g.NP + o]);
    }
  }
}

}  // namespace flow
}  // namespace mfx

using namespace mfx;
using namespace mfx::flow;

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
    nb += H * H + g.NP * H;
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
  // per layer saved state (training): x_in (D), h1, h2 (bf16, H each), raw (NPmax)
  const size_t per_layer = (size_t)n * (D + G.NPmax) * 4 + (size_t)n * H * 2 * 2;
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
      float* xsave = (float*)base;
      A.save_x = xsave;   // the kernel keeps the layer input for the backward pass
      A.save_raw = xsave + (size_t)n * D;
      A.save_h1 = (unsigned short*)(A.save_raw + (size_t)n * G.NPmax);
      A.save_h2 = A.save_h1 + (size_t)n * H;
    }
    const size_t shmem = (size_t)TILE * G.layer[l].NP * 4;
    hipLaunchKernelGGL(k_layer, dim3(nblk), dim3(NT), shmem, s, A, params, (const unsigned short*)params_bf16);
    cur = x_out;
  }
  return hip_status(hipGetLastError());
}

int64_t mfx_flow_save_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G)) return -1;
  return (int64_t)L * ((int64_t)n * (D + G.NPmax) * 4 + (int64_t)n * H * 2 * 2);
}

int64_t mfx_flow_partial_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G)) return -1;
  int maxlp = 0;
  for (int l = 0; l < L; ++l) {
    const int lp = G.layer[l].ob3 + G.layer[l].NP - G.layer[l].oW1;
    maxlp = lp > maxlp ? lp : maxlp;
  }
  return ((n + TILE - 1) / TILE) * (int64_t)maxlp * 4;
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
  const size_t per_layer = (size_t)n * (D + G.NPmax) * 4 + (size_t)n * H * 2 * 2;
  for (int l = L - 1; l >= 0; --l) {
    const char* base = (const char*)save + per_layer * l;
    BwdArgs A;
    A.geo = G;
    A.layer = l;
    A.n = n;
    A.x_in = (const float*)base;
    A.raw = A.x_in + (size_t)n * D;
    A.h1 = (const unsigned short*)(A.raw + (size_t)n * G.NPmax);
    A.h2 = A.h1 + (size_t)n * H;
    A.gl = gl;
    A.gx = gx;
    A.partial = partial;
    const size_t shmem = (size_t)TILE * G.layer[l].NP * 4;
    hipLaunchKernelGGL(k_layer_bwd, dim3(nblk), dim3(NT), shmem, s, A, params, (const unsigned short*)params_bf16);
    const int nlp = G.layer[l].ob3 + G.layer[l].NP - G.layer[l].oW1;
    hipLaunchKernelGGL(k_reduce_grad, dim3((nlp + 63) / 64), dim3(256), 0, s, partial, nblk, nlp,
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
