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
// tiles, so that a small training batch still puts two waves on every SIMD) and take the
// tile through every layer in one launch.  The 128x128 and 128xNP GEMMs run on bf16 MFMA
// (v_mfma_f32_16x16x16bf16, fp32 accumulation, bf16 weights), dW1 on fp32 MFMA
// (v_mfma_f32_16x16x4f32); the spline math is per-lane VALU.  Weight-gradient partials
// are written per split of tiles and reduced in a fixed order.  A training step is five
// launches: the forward pass with the minibatch draw, the per-sample seed, the data and
// weight gradients, their reduction, and Adam with the refresh of the bf16 weights.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "mfx_api_common.h"
#include "mfx_rng.h"
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
// bin widths / heights from raw logits: softmax with a minimum value minv per bin
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

// Bin of v along the cumulative sums c[j] = w[0] + ... + w[j-1] (c[0] = 0, c[K] = 1, summed
// in that order): k = the number of j in 1..K-1 with v >= c[j] (the sums increase, so c[k] <=
// v < c[k+1]), with the edges c[k] and c[k+1].  The sums are formed on the fly, one at a time,
// rather than kept in an array.
struct Edge {
  int k;
  float lo, hi;
};
__device__ __forceinline__ Edge bin_of(const float* w, float v) {
  Edge e{0, 0.f, w[0]};
  float c = 0.f;
#pragma unroll
  for (int j = 1; j < K; ++j) {
    c = c + w[j - 1];                                  // c[j]
    const float cn = j == K - 1 ? 1.f : c + w[j];      // c[j + 1]
    if (v >= c) {
      e.k = j;
      e.lo = c;
      e.hi = cn;
    }
  }
  return e;
}
// the edges c[k], c[k + 1] of bin k of the same cumulative sums
__device__ __forceinline__ Edge edges_at(const float* w, int k) {
  Edge e{k, 0.f, w[0]};
  float c = 0.f;
#pragma unroll
  for (int j = 1; j < K; ++j) {
    c = c + w[j - 1];
    const float cn = j == K - 1 ? 1.f : c + w[j];
    if (j == k) {
      e.lo = c;
      e.hi = cn;
    }
  }
  return e;
}

struct SplineEval {
  float z, ld;  // output and log derivative
};

__device__ __forceinline__ SplineEval rq_forward(const float* raw, float x) {
  float w[K], h[K];
  softmax_bins(raw, MINW, w);
  softmax_bins(raw + K, MINH, h);
  x = fminf(fmaxf(x, 0.f), 1.f);
  const Edge ex = bin_of(w, x);
  const int k = ex.k;
  const Edge ey = edges_at(h, k);
  const float a = MIND + softplus(raw[2 * K + k]), c = MIND + softplus(raw[2 * K + k + 1]);
  const float wk = ex.hi - ex.lo, hk = ey.hi - ey.lo;
  const float xi = fminf(fmaxf((x - ex.lo) / wk, 0.f), 1.f);
  const float s = hk / wk;
  const float t = xi * (1.f - xi);
  const float num = hk * (s * xi * xi + a * t);
  const float den = s + (a + c - 2.f * s) * t;
  const float P = c * xi * xi + 2.f * s * t + a * (1.f - xi) * (1.f - xi);
  SplineEval r;
  r.z = ey.lo + num / den;
  r.ld = 2.f * __logf(s) + __logf(P) - 2.f * __logf(den);
  return r;
}

__device__ __forceinline__ SplineEval rq_inverse(const float* raw, float z) {
  float w[K], h[K];
  softmax_bins(raw, MINW, w);
  softmax_bins(raw + K, MINH, h);
  z = fminf(fmaxf(z, 0.f), 1.f);
  const Edge ey = bin_of(h, z);
  const int k = ey.k;
  const Edge ex = edges_at(w, k);
  const float a = MIND + softplus(raw[2 * K + k]), c = MIND + softplus(raw[2 * K + k + 1]);
  const float wk = ex.hi - ex.lo, hk = ey.hi - ey.lo;
  const float s = hk / wk;
  const float d = z - ey.lo;
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
  r.z = ex.lo + xi * wk;                                          // x
  r.ld = -(2.f * __logf(s) + __logf(P) - 2.f * __logf(den));     // log|dx/du|
  return r;
}

// Gradient of gz*z + gl*ld w.r.t. the raw parameters (written to graw[NPD]) and x.
__device__ __forceinline__ float rq_backward(const float* raw, float x, float gz, float gl, float* graw) {
  float w[K], h[K];
  softmax_bins(raw, MINW, w);
  softmax_bins(raw + K, MINH, h);
  const bool inside = x > 0.f && x < 1.f;
  x = fminf(fmaxf(x, 0.f), 1.f);
  const Edge ex = bin_of(w, x);
  const int k = ex.k;
  const Edge ey = edges_at(h, k);
  const float ra = raw[2 * K + k], rc = raw[2 * K + k + 1];
  const float a = MIND + softplus(ra), c = MIND + softplus(rc);
  const float wk = ex.hi - ex.lo, hk = ey.hi - ey.lo;
  const float xi = fminf(fmaxf((x - ex.lo) / wk, 0.f), 1.f);
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
  // chain to bin quantities: d/d w_j and d/d h_j are GX (GY) below bin k, Gw (Ghk) at k
  const float GX = -Gxi / wk;
  const float Gw = -Gxi * xi / wk - Gs * s / wk;
  const float Ghk = Gh + Gs / wk;
  // through the softmax: w = min + (1 - K min) sm
  float sw = 0.f, sh = 0.f;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    sw += (w[j] - MINW) * ((j < k ? GX : 0.f) + (j == k ? Gw : 0.f));
    sh += (h[j] - MINH) * ((j < k ? GY : 0.f) + (j == k ? Ghk : 0.f));
  }
  const float nw = 1.f - K * MINW, nh = 1.f - K * MINH;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    graw[j] = (w[j] - MINW) * (((j < k ? GX : 0.f) + (j == k ? Gw : 0.f)) - sw / nw);
    graw[K + j] = (h[j] - MINH) * (((j < k ? GY : 0.f) + (j == k ? Ghk : 0.f)) - sh / nh);
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

// ------------------------------------------------------------------ whole-flow kernels
// Every kernel below takes its 16-sample tile (or, for the weight gradients, a split of
// tiles) through all L layers in one launch.  The coupling layers act on each sample
// separately, so a tile's state between layers (coordinates, log-det, incoming gradient)
// stays in LDS, and only what the backward pass needs goes to memory.  The layers' weights
// are MFMA B operands read straight from the bf16 copies (8 layers of 84 KB stay in L2).
// NBM bounds the transformed set of every layer (it sizes the LDS arrays); each layer's
// own geometry is read at run time.
constexpr int WP = H + 8;       // padded row pitch (bf16 elements) of h1/h2 in LDS: 16-byte
                                // rows, and the 16 rows of an MFMA operand in distinct banks

// Saved state of one layer for the backward pass, at a fixed per-layer stride: the layer
// input x (n, D) and the spline parameters raw (n, NPmax) in fp32, then the bf16 activations
// h1 and h2 tile-transposed ([tile][column][sample]) over whole tiles.
struct SaveLayout {
  int64_t raw, h1, h2, bytes;   // byte offsets from the layer's base, and the stride
};
// The bf16 deltas of one layer, tile-transposed over whole tiles: go [nt16][NP], gh2 and
// gh1 [nt16][H], at a fixed per-layer stride.
struct DeltaLayout {
  int64_t o, h2, h1, bytes;
};

__device__ __forceinline__ bf16x4 ld_bf16x4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
typedef short bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 ld_bf16x8(const unsigned short* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}
// acc += A B over 32 consecutive k: the lanes of quarter g hold k = K0 + 8 g .. K0 + 8 g + 7 of
// their row (A) and column (B) in one 16-byte load each, and two MFMAs take the lower and
// upper four (the k order inside the sum is a fixed permutation, the same for A and B).
__device__ __forceinline__ f32x4 mfma_k32(bf16x8 a, bf16x8 b, f32x4 acc) {
  acc = mfma_bf16(__builtin_shufflevector(a, a, 0, 1, 2, 3), __builtin_shufflevector(b, b, 0, 1, 2, 3), acc);
  return mfma_bf16(__builtin_shufflevector(a, a, 4, 5, 6, 7), __builtin_shufflevector(b, b, 4, 5, 6, 7), acc);
}

// first index k of the normalised running sum with cdf[k] >= u
__device__ __forceinline__ int64_t cdf_search(const double* __restrict__ cdf, int64_t n, double u) {
  int64_t lo = 0, hi = n - 1;          // invariant: the answer lies in [lo, hi]
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The same index from a guide table: guide[m] = the first k with cdf[k] >= m / M, so the
// answer for u lies at or after guide[floor(u M)] (with that bucket's threshold at most
// u); usually it is that entry, else a doubling search from it brackets it for a
// bisection.  Two or three dependent loads instead of log2(n).
__device__ __forceinline__ int64_t cdf_search_guided(const double* __restrict__ cdf, int64_t n,
                                                     const int* __restrict__ guide, int64_t M, double u) {
  int64_t m = (int64_t)(u * (double)M);
  if (m >= M) m = M - 1;
  while (m > 0 && (double)m / (double)M > u) --m;
  int64_t lo = guide[m];
  if (lo >= n - 1) return n - 1;
  if (cdf[lo] >= u) return lo;
  // cdf[lo] < u: the answer lies in (lo, n - 1]
  int64_t step = 1, hi = lo + 1;
  while (hi < n - 1 && cdf[hi] < u) {
    lo = hi;
    step *= 2;
    hi = lo + step < n - 1 ? lo + step : n - 1;
  }
  lo = lo + 1;                                 // answer in [lo, hi]
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

struct FlowArgs {
  FlowGeom geo;
  int64_t n;
  const float* x_in;       // (n, D); null when the batch is drawn from a pool
  // minibatch draw: sample e takes pool point k, the first with cdf[k] >= u_e, where u_e
  // is the first 53-bit Philox uniform of counter (e, 0, 0, draw) under key (seed, 0x6d66)
  const double* cdf;
  int64_t npool;
  const int* guide;        // optional guide table of cdf (nguide buckets, mfx_flow_cdf_guide)
  int64_t nguide;
  const float* pool_y;     // (npool, D)
  const double* pool_f;    // optional (npool): integrand values
  uint32_t seed, draw;
  float* y_out;            // optional (n, D): the drawn points
  double* f_out;           // optional (n): their integrand values
  float* x_out;            // optional (n, D): the flow's output
  float* logdet;           // (n): sum over the layers of the log-dets (written)
  char* save;              // optional: saved state of layer l at save + l * sl.bytes
  SaveLayout sl;
  int inverse;
};

// Register budgets: at NBM = 4 (the 8-dim flow) that of 4 waves per SIMD, two 8-wave tiles
// per CU, which the training batch needs to fill the chip; above, that of 3 waves.
template <int NBM>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NBM <= 4 ? 4 : 3)))
k_flow(FlowArgs A, const float* __restrict__ P,
                                             const unsigned short* __restrict__ Pb) {
  constexpr int NPM = ((NBM * NPD + 15) / 16) * 16;
  const int D = A.geo.D, L = A.geo.L;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4, k8 = (lane >> 4) * 8;
  __shared__ float xs[TILE * DMAX], xa[TILE * DMAX], lds_ld[TILE * DMAX], ldacc[TILE];
  __shared__ __attribute__((aligned(16))) unsigned short h1s[TILE * WP];
  __shared__ __attribute__((aligned(16))) unsigned short h2s[TILE * WP];
  __shared__ float raws[TILE * NPM];
  __shared__ int64_t pick[TILE];
  const int64_t s0 = (int64_t)blockIdx.x * TILE;
  if (tid < TILE) {
    const int64_t e = s0 + tid;
    int64_t k = 0;
    if (A.cdf && e < A.n) {
      double u0, u1;
      philox_pair((uint64_t)e, 0u, A.seed, A.draw, u0, u1);
      k = A.guide ? cdf_search_guided(A.cdf, A.npool, A.guide, A.nguide, u0) : cdf_search(A.cdf, A.npool, u0);
      if (A.f_out) A.f_out[e] = A.pool_f[k];
    }
    pick[tid] = k;
    ldacc[tid] = 0.f;
  }
  __syncthreads();
  for (int q = tid; q < TILE * DMAX; q += NT) {
    const int s = q / DMAX, d = q % DMAX;
    const int64_t e = s0 + s;
    const bool ok = e < A.n && d < D;
    const float v = ok ? (A.cdf ? A.pool_y[pick[s] * D + d] : A.x_in[e * D + d]) : 0.5f;
    xs[q] = v;
    if (ok && A.y_out) A.y_out[e * D + d] = v;
  }
  for (int step = 0; step < L; ++step) {
    const int l = A.inverse ? L - 1 - step : step;
    const LayerGeom& g = A.geo.layer[l];
    const int nB = g.nB, NP = g.NP;
    char* base = A.save ? A.save + A.sl.bytes * l : nullptr;
    // the weight operands are requested a phase ahead of their use, so that the loads (L2
    // hits) are in flight while the earlier phases run: here the W1 column and bias of
    // layer 1 and the W2 fragments and bias of this wave's column tile; the W3 fragments
    // and biases of its (up to two) column tiles of the output layer before the W2 GEMM
    static_assert(NT == 4 * H, "four samples per thread in layer 1");
    constexpr int NU3 = (NPM / 16 + WPB - 1) / WPB;
    constexpr int NUP = NU3 < 2 ? NU3 : 2;   // W3 tiles whose operands are prefetched
    const int j1 = tid & (H - 1);
    float w1[DMAX];
#pragma unroll
    for (int i = 0; i < DMAX; ++i) w1[i] = i < g.nA ? P[g.oW1 + i * H + j1] : 0.f;
    const float b1 = P[g.ob1 + j1];
    const float b2 = P[g.ob2 + 16 * wid + r];
    // the coordinate maps this thread reads with run-time indices below
    const int ai = g.A[tid % DMAX];
    const int bi = g.B[tid % nB];
    bf16x8 w2f[H / 32];
#pragma unroll
    for (int c = 0; c < H / 32; ++c) w2f[c] = ld_bf16x8(Pb + g.oW2t + (16 * wid + r) * H + 32 * c + k8);
    __syncthreads();                       // xs holds this layer's input
    static_assert(TILE * DMAX <= NT, "one input coordinate per thread");
    for (int q = tid; q < TILE * DMAX; q += NT) {
      const int s = q / DMAX, i = q % DMAX;
      xa[q] = i < g.nA ? xs[s * DMAX + ai] : 0.f;
      const int64_t e = s0 + s;
      if (base && e < A.n && i < D) reinterpret_cast<float*>(base)[e * D + i] = xs[q];
    }
    __syncthreads();
    // layer 1 on VALU: 16 x 128 outputs; thread t computes column t % H for the samples
    // t / H + 4 u
    {
      const int j = j1, sb = tid / H;
      const float bj = b1;
#pragma unroll
      for (int u = 0; u < TILE / 4; ++u) {
        const int s = sb + 4 * u;
        float v = bj;
#pragma unroll
        for (int i = 0; i < DMAX; ++i)
          if (i < g.nA) v += xa[s * DMAX + i] * w1[i];
        h1s[s * WP + j] = f2bf(fmaxf(v, 0.f));
      }
    }
    __syncthreads();
    // layer 2: h2 = relu(h1 W2 + b2), B operand W2t [out][in]; wave w: column tile 16 w
    bf16x8 w3f[NUP][H / 32];
    float b3[NUP];
#pragma unroll
    for (int u = 0; u < NUP; ++u) {
      const int n0 = 16 * (wid + WPB * u);
      b3[u] = n0 < NP ? P[g.ob3 + n0 + r] : 0.f;
#pragma unroll
      for (int c = 0; c < H / 32; ++c)
        w3f[u][c] = n0 < NP ? ld_bf16x8(Pb + g.oW3t + (n0 + r) * H + 32 * c + k8) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
    {
      const int n0 = 16 * wid;
      const float bias = b2;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int K0 = 0; K0 < H; K0 += 32) acc = mfma_k32(ld_bf16x8(h1s + r * WP + K0 + k8), w2f[K0 / 32], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = kq + i, col = n0 + r;
        h2s[row * WP + col] = f2bf(fmaxf(acc[i] + bias, 0.f));
      }
    }
    __syncthreads();
    // layer 3: raw = h2 W3 + b3, B operand W3t [NP][H]; wave w: column tiles w, w + 8, ...
#pragma unroll
    for (int u = 0; u < NU3; ++u) {
      const int n0 = 16 * (wid + WPB * u);
      if (n0 < NP) {
        const float bias = u < NUP ? b3[u < NUP ? u : 0] : P[g.ob3 + n0 + r];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int K0 = 0; K0 < H; K0 += 32)
          acc = mfma_k32(ld_bf16x8(h2s + r * WP + K0 + k8),
                         u < NUP ? w3f[u < NUP ? u : 0][K0 / 32] : ld_bf16x8(Pb + g.oW3t + (n0 + r) * H + K0 + k8), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = kq + i, col = n0 + r;
          raws[row * NPM + col] = acc[i] + bias;
        }
      }
    }
    __syncthreads();
    // spline on every (sample, transformed dimension); the log-dets of one sample are
    // summed by one thread afterwards in a fixed order
    for (int q = tid; q < TILE * nB; q += NT) {             // one item per thread (TILE nB <= NT)
      const int s = q / nB, b = q % nB;
      const int d = bi;
      const float* raw = raws + s * NPM + b * NPD;
      const SplineEval sp = A.inverse ? rq_inverse(raw, xs[s * DMAX + d]) : rq_forward(raw, xs[s * DMAX + d]);
      xs[s * DMAX + d] = sp.z;
      lds_ld[s * DMAX + b] = sp.ld;
    }
    __syncthreads();
    if (tid < TILE) {
      float ld = 0.f;
      for (int b = 0; b < nB; ++b) ld += lds_ld[tid * DMAX + b];
      ldacc[tid] += ld;
    }
    if (base) {
      unsigned short* sh1 = reinterpret_cast<unsigned short*>(base + A.sl.h1);
      unsigned short* sh2 = reinterpret_cast<unsigned short*>(base + A.sl.h2);
      float* sraw = reinterpret_cast<float*>(base + A.sl.raw);
#pragma unroll
      for (int u = 0; u < TILE * H / NT; ++u) {
        const int q = tid + u * NT;
        const int j = q / TILE, s = q % TILE;
        const bool ok = s0 + s < A.n;
        sh1[s0 * H + q] = ok ? h1s[s * WP + j] : (unsigned short)0;
        sh2[s0 * H + q] = ok ? h2s[s * WP + j] : (unsigned short)0;
      }
#pragma unroll
      for (int u = 0; u < (TILE * NPM + NT - 1) / NT; ++u) {
        const int q = tid + u * NT;
        if (q < TILE * NP) {
          const int s = q / NP, c = q % NP;
          const int64_t e = s0 + s;
          if (e < A.n) sraw[e * A.geo.NPmax + c] = raws[s * NPM + c];
        }
      }
    }
  }
  __syncthreads();
  if (tid < TILE && s0 + tid < A.n) A.logdet[s0 + tid] = ldacc[tid];
  if (A.x_out)
    for (int q = tid; q < TILE * D; q += NT) {
      const int s = q / D, d = q % D;
      const int64_t e = s0 + s;
      if (e < A.n) A.x_out[e * D + d] = xs[s * DMAX + d];
    }
}

// ------------------------------------------------------------------ backward
// Two kernels, both on the bf16 matrix cores with fp32 accumulation: the activations and
// weights are the bf16 values of the forward pass, the deltas are rounded to bf16 like them.
//  * k_flow_bwd: per 16-sample tile, the layers in reverse: the spline backward, the
//    input-gradient products gh2 = (go W3^T) relu'(h2) and gh1 = (gh2 W2^T) relu'(h1), the
//    input gradient (carried to the next layer in LDS), and the deltas go, gh2, gh1 written
//    tile-transposed for the weight gradients;
//  * k_flow_wgrad: the weight gradients of every layer as GEMMs with the samples as K, one
//    partial gradient per split of tiles, reduced in a fixed order (k_reduce_grad).
struct BwdArgs {
  FlowGeom geo;
  int64_t n;
  const char* save;        // the forward pass's saved state
  SaveLayout sl;
  char* delta;             // layer l's deltas at delta + l * dl.bytes
  DeltaLayout dl;
  const float* gl;         // (n): d objective / d log q, per sample
  const float* gx_in;      // optional (n, D): d objective / d flow output (null: zero)
  float* gx_out;           // optional (n, D): d objective / d flow input
  float* partial;          // (nsplit, nparams): weight-gradient partials
};

// The data-gradient kernel runs 4 waves per tile (each wave two column tiles of gh2 / gh1).
// For flows of up to 4 transformed dimensions per layer its register budget is that of 2
// waves per SIMD, two tiles per CU (the training batch of 8,192 samples is 512 tiles, two
// per CU), which leaves room for the GEMMs' weight operands to be requested ahead of their
// use; above that, 3 waves per SIMD.
constexpr int WPB_B = 4;
constexpr int NT_B = 64 * WPB_B;

template <int NBM>
__global__ void __launch_bounds__(NT_B) __attribute__((amdgpu_waves_per_eu(NBM <= 4 ? 2 : 3))) k_flow_bwd(BwdArgs A, const float* __restrict__ P,
                                                   const unsigned short* __restrict__ Pb) {
  constexpr int NPM = ((NBM * NPD + 15) / 16) * 16;
  constexpr int NPP = NPM + 8;                 // padded bf16 row of go (16-byte rows, distinct banks)
  static_assert(H / 16 == 2 * WPB_B, "two column tiles of gh2 / gh1 per wave");
  static_assert(TILE * DMAX == NT_B, "one coordinate per thread");
  const int D = A.geo.D, L = A.geo.L;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;   // bf16 MFMA: four consecutive k per lane
  const int k8 = (lane >> 4) * 8;                  // ... eight in the 16-byte loads of mfma_k32
  const int cw = wid;                              // this wave's column tiles: cw and cw + 4
  const int64_t s0 = (int64_t)blockIdx.x * TILE;
  __shared__ float xs[TILE * DMAX], gxs[TILE * DMAX];
  __shared__ __attribute__((aligned(16))) unsigned short h1b[TILE * WP], h2b[TILE * WP], ghb[TILE * WP];
  __shared__ __attribute__((aligned(16))) unsigned short gob[TILE * NPP];
  // go: the tile's spline parameters, overwritten in place by their gradient
  __shared__ __attribute__((aligned(16))) float go[TILE * NPM], gh1[TILE * H], W1s[DMAX * H];
  // the layer's coordinate maps (read per lane with run-time indices, so staged in LDS with
  // the layer's state rather than loaded from the argument block where they are used) and
  // the tile's d objective / d log q
  __shared__ int Al[DMAX], Bl[DMAX];
  __shared__ float gls[TILE];
  {
    const int s = tid / DMAX, d = tid % DMAX;
    const int64_t e = s0 + s;
    gxs[tid] = (A.gx_in && e < A.n && d < D) ? A.gx_in[e * D + d] : 0.f;
    if (tid < TILE) gls[tid] = s0 + tid < A.n ? A.gl[s0 + tid] : 0.f;
  }
  for (int l = L - 1; l >= 0; --l) {
    const LayerGeom& g = A.geo.layer[l];
    const int nB = g.nB, NP = g.NP;
    const char* base = A.save + A.sl.bytes * l;
    const float* sx = reinterpret_cast<const float*>(base);
    const float* sraw = reinterpret_cast<const float*>(base + A.sl.raw);
    const unsigned short* sh1 = reinterpret_cast<const unsigned short*>(base + A.sl.h1);
    const unsigned short* sh2 = reinterpret_cast<const unsigned short*>(base + A.sl.h2);
    char* dbase = A.delta + A.dl.bytes * l;
    unsigned short* d_o = reinterpret_cast<unsigned short*>(dbase + A.dl.o);
    unsigned short* d_h2 = reinterpret_cast<unsigned short*>(dbase + A.dl.h2);
    unsigned short* d_h1 = reinterpret_cast<unsigned short*>(dbase + A.dl.h1);
    __syncthreads();                             // the previous layer is done with the LDS
    // the loads of this layer's state, unrolled to their bounds and made unconditional
    // (out-of-range lanes read an in-range address and drop the value) so that they are
    // all in flight together
    {
      const int s = tid / DMAX, d = tid % DMAX;
      const int64_t e = s0 + s;
      xs[tid] = (e < A.n && d < D) ? sx[e * D + d] : 0.5f;
      if (tid < DMAX) {
        Al[tid] = g.A[tid];
        Bl[tid] = g.B[tid];
      }
    }
    {
      const int nq = g.nA * H / 4;             // float4s of W1 (g.oW1 is a multiple of 4; nA >= 1)
      float4 v[DMAX * H / 4 / NT_B];
#pragma unroll
      for (int u = 0; u < DMAX * H / 4 / NT_B; ++u) {
        const int q = tid + u * NT_B;
        v[u] = *reinterpret_cast<const float4*>(P + g.oW1 + 4 * (q < nq ? q : nq - 1));
      }
#pragma unroll
      for (int u = 0; u < DMAX * H / 4 / NT_B; ++u) {
        const int q = tid + u * NT_B;
        if (q < nq) *reinterpret_cast<float4*>(W1s + 4 * q) = v[u];
      }
    }
#pragma unroll
    for (int u = 0; u < TILE * H / 2 / NT_B; ++u) {
      const int q = tid + u * NT_B;            // samples s, s + 1 of column j: one 4-byte load
      const int j = (2 * q) / TILE, s = (2 * q) % TILE;   // saved activations are tile-transposed
      const unsigned v1 = *reinterpret_cast<const unsigned*>(sh1 + s0 * H + 2 * q);
      const unsigned v2 = *reinterpret_cast<const unsigned*>(sh2 + s0 * H + 2 * q);
      h1b[s * WP + j] = (unsigned short)v1;
      h1b[(s + 1) * WP + j] = (unsigned short)(v1 >> 16);
      h2b[s * WP + j] = (unsigned short)v2;
      h2b[(s + 1) * WP + j] = (unsigned short)(v2 >> 16);
    }
    // spline parameters of the transformed dimensions, four at a time; the padding columns
    // (and samples past n) are zero and stay zero in the gradient
    const int nraw = nB * NPD, np4 = NP / 4;
    {
      constexpr int NR = (TILE * NPM / 4 + NT_B - 1) / NT_B;
      float4 v[NR];
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const int q0 = tid + u * NT_B, q = q0 < TILE * np4 ? q0 : TILE * np4 - 1;
        const int s = q / np4, c = 4 * (q % np4);
        const int64_t e = s0 + s < A.n ? s0 + s : A.n - 1;
        v[u] = *reinterpret_cast<const float4*>(sraw + e * A.geo.NPmax + c);
      }
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const int q = tid + u * NT_B;
        if (q < TILE * np4) {
          const int s = q / np4, c = 4 * (q % np4);
          float4 w = s0 + s < A.n ? v[u] : float4{0.f, 0.f, 0.f, 0.f};
          if (c + 0 >= nraw) w.x = 0.f;
          if (c + 1 >= nraw) w.y = 0.f;
          if (c + 2 >= nraw) w.z = 0.f;
          if (c + 3 >= nraw) w.w = 0.f;
          *reinterpret_cast<float4*>(go + s * NPM + c) = w;
        }
      }
    }
    __syncthreads();
    // spline backward: d raw (in place in go) and d x_B (replaces the incoming gradient of
    // x_B); each item reads its parameters before it writes their gradient
    for (int q = tid; q < TILE * nB; q += NT_B) {
      const int s = q / nB, b = q % nB;
      const int64_t e = s0 + s;
      if (e >= A.n) continue;
      const int d = Bl[b];
      float* gr = go + s * NPM + b * NPD;
      gxs[s * DMAX + d] = rq_backward(gr, xs[s * DMAX + d], gxs[s * DMAX + d], gls[s], gr);
    }
    __syncthreads();
    // go in bf16: the A operand of gh2 (LDS) and the delta of dW3 / b3 (tile-transposed)
    for (int q = tid; q < TILE * NP; q += NT_B) {
      const int c = q / TILE, s = q % TILE;
      const unsigned short v = f2bf(go[s * NPM + c]);
      gob[s * NPP + c] = v;
      d_o[s0 * NP + q] = v;
    }
    __syncthreads();
    bf16x8 w2f[2][H / 32];                       // gh1's B operands, requested in the gh2 phase
    // gh2 = (go W3^T) * relu'(h2): B[k = np][n = h] = W3[h][np], the row h of W3n; wave w
    // computes the column tiles cw and cw + 4 (one A operand, two B operands per k step)
    {
      f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
      const unsigned short* w3[2] = {Pb + g.oW3n + (16 * cw + r) * NP, Pb + g.oW3n + (16 * (cw + WPB_B) + r) * NP};
      // the B operands of every k step are requested before the products (a step past NP
      // reads an in-range address and is skipped), and the next GEMM's W2 operands with them
      constexpr bool PF3 = NPM / 32 <= 6;          // W3 prefetched up to the 8-dim flow's width
      bf16x8 w3f[2][PF3 ? NPM / 32 : 1];
      if constexpr (PF3) {
#pragma unroll
        for (int K0 = 0; K0 + 32 <= NPM; K0 += 32)
#pragma unroll
          for (int u = 0; u < 2; ++u) w3f[u][K0 / 32] = ld_bf16x8(w3[u] + (K0 + 32 <= NP ? K0 : 0) + k8);
      }
#pragma unroll
      for (int K0 = 0; K0 < H; K0 += 32)
#pragma unroll
        for (int u = 0; u < 2; ++u) w2f[u][K0 / 32] = ld_bf16x8(Pb + g.oW2n + (16 * (cw + WPB_B * u) + r) * H + K0 + k8);
#pragma unroll
      for (int K0 = 0; K0 + 32 <= NPM; K0 += 32)
        if (K0 + 32 <= NP) {
          const bf16x8 a = ld_bf16x8(gob + r * NPP + K0 + k8);
#pragma unroll
          for (int u = 0; u < 2; ++u)
            acc[u] = mfma_k32(a, PF3 ? w3f[u][PF3 ? K0 / 32 : 0] : ld_bf16x8(w3[u] + K0 + k8), acc[u]);
        }
      if (NP % 32) {                               // NP = 32 m + 16: the last 16 k
        const int k0 = NP - 16;
        const bf16x4 a = ld_bf16x4(gob + r * NPP + k0 + kq);
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[u] = mfma_bf16(a, ld_bf16x4(w3[u] + k0 + kq), acc[u]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int n0 = 16 * (cw + WPB_B * u);
        unsigned short out[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = kq + i, col = n0 + r;
          const unsigned short v = f2bf(bf2f(h2b[row * WP + col]) > 0.f ? acc[u][i] : 0.f);
          ghb[row * WP + col] = v;
          out[i] = v;
        }
        // the four samples of column n0 + r: one 8-byte store
        *reinterpret_cast<uint2*>(d_h2 + s0 * H + (n0 + r) * TILE + kq) =
            make_uint2((unsigned)out[0] | ((unsigned)out[1] << 16), (unsigned)out[2] | ((unsigned)out[3] << 16));
      }
    }
    __syncthreads();
    // gh1 = (gh2 W2^T) * relu'(h1): B[k = o][n = i] = W2[i][o], the row i of W2n
    {
      f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int K0 = 0; K0 < H; K0 += 32) {
        const bf16x8 a = ld_bf16x8(ghb + r * WP + K0 + k8);
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[u] = mfma_k32(a, w2f[u][K0 / 32], acc[u]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int n0 = 16 * (cw + WPB_B * u);
        unsigned short out[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = kq + i, col = n0 + r;
          const float v = bf2f(h1b[row * WP + col]) > 0.f ? acc[u][i] : 0.f;
          gh1[row * H + col] = v;
          out[i] = f2bf(v);
        }
        *reinterpret_cast<uint2*>(d_h1 + s0 * H + (n0 + r) * TILE + kq) =
            make_uint2((unsigned)out[0] | ((unsigned)out[1] << 16), (unsigned)out[2] | ((unsigned)out[3] << 16));
      }
    }
    __syncthreads();
    // d x_A += gh1 W1^T (fp32): 8 lanes per (sample, input), 16 terms each, summed by
    // shuffles within aligned groups of 8 lanes (the summation order is fixed)
    static_assert((NT_B / 8) % DMAX == 0, "a thread's input i is the same in every pass");
    const int iA = (tid >> 3) % DMAX;
    const int dA = Al[iA];
    for (int q0 = 0; q0 < TILE * DMAX * 8; q0 += NT_B) {
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
      if (part == 0 && i < g.nA) gxs[s * DMAX + dA] += v;
    }
  }
  __syncthreads();
  if (A.gx_out && tid < TILE * D) {
    const int s = tid / D, d = tid % D;
    const int64_t e = s0 + s;
    if (e < A.n) A.gx_out[e * D + d] = gxs[s * DMAX + d];
  }
}

// Weight gradients: one block per (split of WG_ST tiles, layer).  The split's saved
// activations h1, h2 (the A operands, shared by every wave) and inputs x_A are staged in LDS
// once; wave w owns whole column tiles (dW3 tiles w, w + 8, ...; dW2, b2, b1 and dW1 tile w)
// and streams its deltas (the B operands) from memory, so every byte of the split is read
// from memory once.  Two tiles form one 32-sample k step (mfma_k32); a lone last tile
// takes one 16-sample step.
constexpr int WG_ST = 16;

__global__ void __launch_bounds__(NT) k_flow_wgrad(BwdArgs A) {
  static_assert(DMAX == 16 && H / 16 == WPB, "dW1 is one 16-row tile; one dW2 column tile per wave");
  const int l = blockIdx.y;
  const LayerGeom& g = A.geo.layer[l];
  const int NP = g.NP, NT3 = NP / 16;
  const int D = A.geo.D;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4, kf = lane >> 4;
  // pair k step: lane quarter g takes samples 8 (g % 2) .. + 7 of tile t + g / 2
  const int ph = (lane >> 5), po = ((lane >> 4) & 1) * 8;
  const char* base = A.save + A.sl.bytes * l;
  const float* sx = reinterpret_cast<const float*>(base);
  const unsigned short* sh1 = reinterpret_cast<const unsigned short*>(base + A.sl.h1);
  const unsigned short* sh2 = reinterpret_cast<const unsigned short*>(base + A.sl.h2);
  const char* dbase = A.delta + A.dl.bytes * l;
  const unsigned short* d_o = reinterpret_cast<const unsigned short*>(dbase + A.dl.o);
  const unsigned short* d_h2 = reinterpret_cast<const unsigned short*>(dbase + A.dl.h2);
  const unsigned short* d_h1 = reinterpret_cast<const unsigned short*>(dbase + A.dl.h1);
  const int64_t ntiles = (A.n + TILE - 1) / TILE;
  const int64_t t0 = (int64_t)blockIdx.x * WG_ST;
  const int nt = (int)(t0 + WG_ST < ntiles ? WG_ST : ntiles - t0);   // tiles of this split
  float* part = A.partial + (size_t)blockIdx.x * A.geo.nparams + g.oW1;
  __shared__ __attribute__((aligned(16))) unsigned short h1l[WG_ST * TILE * H], h2l[WG_ST * TILE * H];
  __shared__ float xl[WG_ST * TILE * DMAX];
  // stage the split's activations ([tile][column][sample], as saved: samples past n are
  // zero there and in the deltas) and its conditioning inputs (zero past n and for i >= nA)
  for (int q = tid; q < nt * TILE * H / 8; q += NT) {
    *reinterpret_cast<uint4*>(h1l + 8 * q) = *reinterpret_cast<const uint4*>(sh1 + t0 * TILE * H + 8 * q);
    *reinterpret_cast<uint4*>(h2l + 8 * q) = *reinterpret_cast<const uint4*>(sh2 + t0 * TILE * H + 8 * q);
  }
  for (int q = tid; q < nt * TILE * DMAX; q += NT) {
    const int i = q % DMAX;
    const int64_t e = t0 * TILE + q / DMAX;
    xl[q] = (i < g.nA && e < A.n) ? sx[e * D + g.A[i]] : 0.f;
  }
  __syncthreads();
  const short one = (short)0x3f80;            // 1.0 in bf16
  const bf16x8 ones8 = {one, one, one, one, one, one, one, one};
  const bf16x4 ones4 = {one, one, one, one};
  const bf16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  const bf16x4 zero4 = {0, 0, 0, 0};
  // ---- dW3 = h2^T go and b3 = ones^T go: column tiles ja = wid + 16 u and jb = ja + 8
  for (int ja = wid; ja < NT3; ja += 2 * WPB) {
    const int jb = ja + WPB;
    const bool hb = jb < NT3;
    f32x4 acc[2][H / 16], bacc[2];
#pragma unroll
    for (int m = 0; m < H / 16; ++m) acc[0][m] = acc[1][m] = f32x4{0.f, 0.f, 0.f, 0.f};
    bacc[0] = bacc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
    int t = 0;
    for (; t + 1 < nt; t += 2) {
      const int64_t tt = t0 + t + ph;
      const bf16x8 b0 = ld_bf16x8(d_o + tt * TILE * NP + (16 * ja + r) * TILE + po);
      const bf16x8 b1 = hb ? ld_bf16x8(d_o + tt * TILE * NP + (16 * jb + r) * TILE + po) : zero8;
#pragma unroll
      for (int m = 0; m < H / 16; ++m) {
        const bf16x8 a = ld_bf16x8(h2l + ((t + ph) * H + 16 * m + r) * TILE + po);
        acc[0][m] = mfma_k32(a, b0, acc[0][m]);
        acc[1][m] = mfma_k32(a, b1, acc[1][m]);
      }
      bacc[0] = mfma_k32(ones8, b0, bacc[0]);
      bacc[1] = mfma_k32(ones8, b1, bacc[1]);
    }
    if (t < nt) {                                // a lone last tile: one 16-sample step
      const int64_t tt = t0 + t;
      const bf16x4 b0 = ld_bf16x4(d_o + tt * TILE * NP + (16 * ja + r) * TILE + kq);
      const bf16x4 b1 = hb ? ld_bf16x4(d_o + tt * TILE * NP + (16 * jb + r) * TILE + kq) : zero4;
#pragma unroll
      for (int m = 0; m < H / 16; ++m) {
        const bf16x4 a = ld_bf16x4(h2l + (t * H + 16 * m + r) * TILE + kq);
        acc[0][m] = mfma_bf16(a, b0, acc[0][m]);
        acc[1][m] = mfma_bf16(a, b1, acc[1][m]);
      }
      bacc[0] = mfma_bf16(ones4, b0, bacc[0]);
      bacc[1] = mfma_bf16(ones4, b1, bacc[1]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = u ? jb : ja;
      if (u && !hb) break;
#pragma unroll
      for (int m = 0; m < H / 16; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) part[g.oW3 - g.oW1 + (16 * m + kq + i) * NP + 16 * j + r] = acc[u][m][i];
      if (kq == 0) part[g.ob3 - g.oW1 + 16 * j + r] = bacc[u][0];
    }
  }
  // ---- dW2 = h1^T gh2 and b2 = ones^T gh2 (column tile wid); b1 = ones^T gh1 and
  // dW1 = x_A^T gh1 (fp32 matrix cores, from the exact inputs) of column tile wid
  {
    const int j = wid;
    f32x4 acc[H / 16], b2acc = {0.f, 0.f, 0.f, 0.f}, b1acc = {0.f, 0.f, 0.f, 0.f}, w1acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < H / 16; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    int t = 0;
    for (; t + 1 < nt; t += 2) {
      const int64_t tt = t0 + t + ph;
      const bf16x8 b = ld_bf16x8(d_h2 + tt * TILE * H + (16 * j + r) * TILE + po);
      const bf16x8 c = ld_bf16x8(d_h1 + tt * TILE * H + (16 * j + r) * TILE + po);
#pragma unroll
      for (int m = 0; m < H / 16; ++m)
        acc[m] = mfma_k32(ld_bf16x8(h1l + ((t + ph) * H + 16 * m + r) * TILE + po), b, acc[m]);
      b2acc = mfma_k32(ones8, b, b2acc);
      b1acc = mfma_k32(ones8, c, b1acc);
    }
    if (t < nt) {
      const int64_t tt = t0 + t;
      const bf16x4 b = ld_bf16x4(d_h2 + tt * TILE * H + (16 * j + r) * TILE + kq);
      const bf16x4 c = ld_bf16x4(d_h1 + tt * TILE * H + (16 * j + r) * TILE + kq);
#pragma unroll
      for (int m = 0; m < H / 16; ++m) acc[m] = mfma_bf16(ld_bf16x4(h1l + (t * H + 16 * m + r) * TILE + kq), b, acc[m]);
      b2acc = mfma_bf16(ones4, b, b2acc);
      b1acc = mfma_bf16(ones4, c, b1acc);
    }
    // dW1: A[i][k] = x_A of sample k (row i < 16), B[k][c] = gh1; four samples per step,
    // lane quarter kf taking samples 4 kf .. 4 kf + 3 of each tile (one 8-byte load of the
    // tile-transposed deltas).  Every tile's load is issued before the products (a tile
    // past the split reads the last one and is skipped), so their latencies overlap.
    {
      uint2 gq[WG_ST];
#pragma unroll
      for (int t = 0; t < WG_ST; ++t) {
        const int tc = t < nt ? t : nt - 1;
        gq[t] = *reinterpret_cast<const uint2*>(d_h1 + (t0 + tc) * TILE * H + (16 * j + r) * TILE + 4 * kf);
      }
#pragma unroll
      for (int t = 0; t < WG_ST; ++t) {
        if (t < nt) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const unsigned wd = u < 2 ? gq[t].x : gq[t].y;
            const float gv = __uint_as_float((u & 1 ? wd >> 16 : wd & 0xffffu) << 16);
            const int sk = t * TILE + 4 * kf + u;
            w1acc = mfma_f32(xl[sk * DMAX + r], gv, w1acc);
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < H / 16; ++m)
#pragma unroll
      for (int i = 0; i < 4; ++i) part[g.oW2 - g.oW1 + (16 * m + kq + i) * H + 16 * j + r] = acc[m][i];
    // fp32 MFMA output: D[4 kf + i][r] in element i
#pragma unroll
    for (int i = 0; i < 4; ++i) part[(4 * kf + i) * H + 16 * j + r] = w1acc[i];
    if (kq == 0) {
      part[g.ob2 - g.oW1 + 16 * j + r] = b2acc[0];
      part[g.ob1 - g.oW1 + 16 * j + r] = b1acc[0];
    }
  }
}

// Sum of the per-split partial gradients: 64 parameters per block, the partials of each
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
  const int64_t k = cdf_search(cdf, n, r[i]);
  for (int d = 0; d < D; ++d) yb[i * D + d] = y[k * D + d];
  if (f) fb[i] = f[k];
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
  // up to SEED_C items per thread (a training batch of up to 8,192 samples) keep their
  // log-ratio and weight in registers, so each log and exp is evaluated once; larger n
  // evaluates them again in the later passes (the same values)
  constexpr int SEED_C = 8;
  const bool cached = n <= SEED_C * SEED_T;
  double lc[SEED_C];
  double mx = -HUGE_VAL;
  if (beta != 0.f) {
    if (cached) {
#pragma unroll
      for (int c = 0; c < SEED_C; ++c) {
        const int i = t + c * SEED_T;
        lc[c] = i < n ? lv(i) : -HUGE_VAL;
        mx = fmax(mx, lc[c]);
      }
    } else {
      for (int i = t; i < n; i += SEED_T) mx = fmax(mx, lv(i));
    }
  }
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
  if (cached) {
#pragma unroll
    for (int c = 0; c < SEED_C; ++c) {
      const int i = t + c * SEED_T;
      lc[c] = i < n ? (w ? w[i] : 1.0) * (beta != 0.f ? exp(lc[c] - mx) : 1.0) : 0.0;   // now v_i
      sum += lc[c];
    }
  } else {
    for (int i = t; i < n; i += SEED_T) sum += (w ? w[i] : 1.0) * (beta != 0.f ? exp(lv(i) - mx) : 1.0);
  }
  red[t] = sum;
  __syncthreads();
  for (int s = SEED_T / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double tot = red[0];
  if (cached) {
#pragma unroll
    for (int c = 0; c < SEED_C; ++c) {
      const int i = t + c * SEED_T;
      if (i < n) gl[i] = (float)(tot > 0.0 ? -lc[c] / tot : 0.0);
    }
  } else {
    for (int i = t; i < n; i += SEED_T) {
      const double v = (w ? w[i] : 1.0) * (beta != 0.f ? exp(lv(i) - mx) : 1.0);
      gl[i] = (float)(tot > 0.0 ? -v / tot : 0.0);
    }
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

// Adam step on the fp32 master parameters of every layer (block row y = layer) together with
// the bf16 copies of the updated hidden/output weights (the same values k_cast writes).
__global__ void k_adam_cast(FlowGeom geo, float* __restrict__ P, const float* __restrict__ grad,
                            float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2, float eps,
                            float bc1, float bc2, float clip, unsigned short* __restrict__ Pb) {
  const LayerGeom& g = geo.layer[blockIdx.y];
  const int nlp = g.ob3 + g.NP - g.oW1;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nlp; q += gridDim.x * blockDim.x) {
    const int i = g.oW1 + q;
    float gi = grad[i];
    if (!isfinite(gi)) gi = 0.f;
    gi = fminf(fmaxf(gi, -clip), clip);
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float p = P[i] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    P[i] = p;
    if (i >= g.oW2 && i < g.ob2) {
      const int qq = i - g.oW2, in = qq / H, o = qq % H;   // W2[in][o]
      Pb[g.oW2n + qq] = f2bf(p);
      Pb[g.oW2t + o * H + in] = f2bf(p);
    } else if (i >= g.oW3 && i < g.ob3) {
      const int qq = i - g.oW3, in = qq / g.NP, o = qq % g.NP;   // W3[in][o]
      Pb[g.oW3n + qq] = f2bf(p);
      Pb[g.oW3t + o * H + in] = f2bf(p);
    }
  }
}

// ------------------------------------------------------------------ training-set and sampling helpers
// Normalised running sum of the training weights, max(w, 0) (a weight that is not a number
// counts as 0), in fp64: per-block inclusive sums over CDF_T * CDF_V consecutive weights
// (each thread sums CDF_V in order, the thread totals are scanned in LDS in a fixed order),
// then the block totals are scanned by one block, then every entry gets its block's offset
// and is divided by the total.  The summation order is fixed, so the result is reproducible.
constexpr int CDF_T = 256, CDF_V = 16, CDF_B = CDF_T * CDF_V;

__device__ __forceinline__ double wclamp(double w) { return w > 0.0 ? w : 0.0; }

__global__ void __launch_bounds__(CDF_T) k_cdf_blocks(const double* __restrict__ w, int64_t n,
                                                      double* __restrict__ cdf, double* __restrict__ btot) {
  __shared__ double ts[CDF_T];
  const int t = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * CDF_B + (int64_t)t * CDF_V;
  double run = 0.0;
  for (int v = 0; v < CDF_V; ++v) {
    const int64_t i = i0 + v;
    if (i < n) {
      run += wclamp(w[i]);
      cdf[i] = run;                     // the thread's own running sum, offset below
    }
  }
  ts[t] = run;
  __syncthreads();
  if (t == 0) {                          // exclusive scan of the thread totals, in order
    double acc = 0.0;
    for (int k = 0; k < CDF_T; ++k) {
      const double v = ts[k];
      ts[k] = acc;
      acc += v;
    }
    btot[blockIdx.x] = acc;
  }
  __syncthreads();
  const double off = ts[t];
  for (int v = 0; v < CDF_V; ++v) {
    const int64_t i = i0 + v;
    if (i < n) cdf[i] += off;
  }
}

// exclusive scan of the block totals in place, in order (one thread); the total goes to
// btot[nb]
__global__ void k_cdf_totals(double* __restrict__ btot, int nb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double v = btot[b];
    btot[b] = acc;
    acc += v;
  }
  btot[nb] = acc;
}

__global__ void k_cdf_finish(double* __restrict__ cdf, int64_t n, const double* __restrict__ btot, int nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tot = btot[nb];
  const double v = cdf[i] + btot[i / CDF_B];
  cdf[i] = tot > 0.0 ? (i == n - 1 ? 1.0 : v / tot) : 0.0;
}

// guide[m] = the first k with cdf[k] >= m / M (n - 1 when there is none), one bisection per
// bucket: the table of cdf_search_guided.
__global__ void k_cdf_guide(const double* __restrict__ cdf, int64_t n, int* __restrict__ guide, int64_t M) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  guide[m] = (int)cdf_search(cdf, n, (double)m / (double)M);
}

// Second moment of defensive-mixture importance weights: sum_i f_i^2 / ((1 - mix) e^{logq_i} + mix),
// in fp64: MOM_G blocks each sum a contiguous range (thread partials in order, then the
// block's threads in order), one thread sums the block partials in order.
constexpr int MOM_G = 256, MOM_T = 256;

__global__ void __launch_bounds__(MOM_T) k_moment(const float* __restrict__ logq, const double* __restrict__ f,
                                                  int64_t n, double mix, double* __restrict__ part) {
  __shared__ double ts[MOM_T];
  const int64_t per = (n + MOM_G - 1) / MOM_G;
  const int64_t b0 = (int64_t)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += MOM_T) {
    const double q = (1.0 - mix) * exp((double)logq[i]) + mix;
    acc += f[i] * f[i] / q;
  }
  ts[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < MOM_T; ++k) s += ts[k];
    part[blockIdx.x] = s;
  }
}

__global__ void k_moment_sum(double* __restrict__ part) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < MOM_G; ++b) s += part[b];
  part[MOM_G] = s;
}

// 1 / q'(y) of the defensive mixture q' = (1 - mix) q + mix, from the log-dets of a forward
// pass (log q = ld) or of an inverse pass (log q = -ld)
__global__ void k_mixture_winv(const float* __restrict__ ld, int64_t n, double sign, double mix,
                               double* __restrict__ winv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) winv[i] = 1.0 / ((1.0 - mix) * exp(sign * (double)ld[i]) + mix);
}

}  // namespace flow
}  // namespace mfx

using namespace mfx;
using namespace mfx::flow;

static int64_t ntiles_of(int64_t n) { return (n + TILE - 1) / TILE; }
static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

static SaveLayout save_layout(const FlowGeom& G, int64_t n) {
  SaveLayout s;
  s.raw = align256(n * G.D * 4);
  s.h1 = s.raw + align256(n * G.NPmax * 4);
  s.h2 = s.h1 + align256(ntiles_of(n) * TILE * H * 2);
  s.bytes = s.h2 + align256(ntiles_of(n) * TILE * H * 2);
  return s;
}

static DeltaLayout delta_layout(const FlowGeom& G, int64_t n) {
  const int64_t nt16 = ntiles_of(n) * TILE;
  DeltaLayout d;
  d.o = 0;
  d.h2 = align256(nt16 * G.NPmax * 2);
  d.h1 = d.h2 + align256(nt16 * H * 2);
  d.bytes = d.h1 + align256(nt16 * H * 2);
  return d;
}

// Weight-gradient splits of WG_ST tiles
static int64_t nsplit_of(int64_t n) { return (ntiles_of(n) + WG_ST - 1) / WG_ST; }

static int make_geom(int D, int L, const int* masks, FlowGeom& G) {
  if (D < 2 || D > DMAX || L < 1 || L > 16) return fail(-2, "flow needs 2 <= D <= 16 and 1 <= L <= 16%s", "");
  if (!masks) return fail(-3, "null masks%s", "");
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

// The kernels are compiled for transformed sets of up to 4, 8 and 16 dimensions (the LDS
// arrays are sized for the bound); a flow runs the smallest that holds its widest layer.
static int nb_bound(const FlowGeom& G) {
  int m = 0;
  for (int l = 0; l < G.L; ++l) m = G.layer[l].nB > m ? G.layer[l].nB : m;
  return m <= 4 ? 4 : (m <= 8 ? 8 : 16);
}

static void launch_flow(const FlowArgs& A, hipStream_t s, const float* P, const unsigned short* Pb) {
  const dim3 grid((unsigned)ntiles_of(A.n));
  switch (nb_bound(A.geo)) {
    case 4: hipLaunchKernelGGL(k_flow<4>, grid, dim3(NT), 0, s, A, P, Pb); break;
    case 8: hipLaunchKernelGGL(k_flow<8>, grid, dim3(NT), 0, s, A, P, Pb); break;
    default: hipLaunchKernelGGL(k_flow<16>, grid, dim3(NT), 0, s, A, P, Pb); break;
  }
}

static FlowArgs flow_args(const FlowGeom& G, int64_t n) {
  FlowArgs A;
  A.geo = G;
  A.n = n;
  A.x_in = nullptr;
  A.cdf = nullptr;
  A.npool = 0;
  A.guide = nullptr;
  A.nguide = 0;
  A.pool_y = nullptr;
  A.pool_f = nullptr;
  A.seed = A.draw = 0;
  A.y_out = nullptr;
  A.f_out = nullptr;
  A.x_out = nullptr;
  A.logdet = nullptr;
  A.save = nullptr;
  A.sl = save_layout(G, n);
  A.inverse = 0;
  return A;
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
  if (!params || !params_bf16) return fail(-3, "null pointer%s", "");
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
  FlowArgs A = flow_args(G, n);
  A.x_in = x_in;
  A.x_out = x_out;
  A.logdet = logdet;
  A.inverse = inverse ? 1 : 0;
  A.save = (save && !inverse) ? (char*)save : nullptr;
  launch_flow(A, (hipStream_t)stream, params, (const unsigned short*)params_bf16);
  return hip_status(hipGetLastError());
}

int mfx_flow_draw_forward(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16,
                          const double* cdf, const int* guide, int64_t nguide, int64_t npool, const float* pool_y,
                          const double* pool_f, uint32_t seed, uint32_t draw, int64_t batch, float* y_out,
                          double* f_out, float* logdet, float* save, void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  if (npool <= 0 || batch < 0) return fail(-2, "npool > 0 and batch >= 0 required%s", "");
  if (guide && nguide <= 0) return fail(-2, "nguide > 0 required with a guide table%s", "");
  if (batch == 0) return 0;
  if (!params || !params_bf16 || !cdf || !pool_y || !logdet || (f_out && !pool_f))
    return fail(-3, "null pointer (pool_f is needed for f_out)%s", "");
  FlowArgs A = flow_args(G, batch);
  A.cdf = cdf;
  A.npool = npool;
  A.guide = guide;
  A.nguide = guide ? nguide : 0;
  A.pool_y = pool_y;
  A.pool_f = pool_f;
  A.seed = seed;
  A.draw = draw;
  A.y_out = y_out;
  A.f_out = f_out;
  A.logdet = logdet;
  A.save = (char*)save;
  launch_flow(A, (hipStream_t)stream, params, (const unsigned short*)params_bf16);
  return hip_status(hipGetLastError());
}

int64_t mfx_flow_save_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G) || n < 0) return -1;
  return (int64_t)L * save_layout(G, n).bytes;
}

int64_t mfx_flow_partial_bytes(int D, int L, const int* masks, int64_t n) {
  FlowGeom G;
  if (make_geom(D, L, masks, G) || n < 0) return -1;
  // the deltas of every layer, then one partial gradient per split
  return (int64_t)L * delta_layout(G, n).bytes + nsplit_of(n) * (int64_t)G.nparams * 4;
}

int mfx_flow_backward(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16, int64_t n,
                      const float* save, const float* gl, float* gx, float* partial, float* grad, void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  if (n <= 0) return n < 0 ? fail(-2, "n must be >= 0%s", "") : 0;
  if (!params || !params_bf16 || !save || !gl || !partial || !grad) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  BwdArgs A;
  A.geo = G;
  A.n = n;
  A.save = (const char*)save;
  A.sl = save_layout(G, n);
  A.delta = (char*)partial;
  A.dl = delta_layout(G, n);
  A.gl = gl;
  A.gx_in = gx;
  A.gx_out = gx;
  A.partial = (float*)((char*)partial + (int64_t)L * A.dl.bytes);
  const int nsplit = (int)nsplit_of(n);
  const dim3 gd((unsigned)ntiles_of(n)), gw((unsigned)nsplit, (unsigned)L);
  switch (nb_bound(G)) {
    case 4:
      hipLaunchKernelGGL(k_flow_bwd<4>, gd, dim3(NT_B), 0, s, A, params, (const unsigned short*)params_bf16);
      break;
    case 8:
      hipLaunchKernelGGL(k_flow_bwd<8>, gd, dim3(NT_B), 0, s, A, params, (const unsigned short*)params_bf16);
      break;
    default:
      hipLaunchKernelGGL(k_flow_bwd<16>, gd, dim3(NT_B), 0, s, A, params, (const unsigned short*)params_bf16);
      break;
  }
  hipLaunchKernelGGL(k_flow_wgrad, gw, dim3(NT), 0, s, A);
  hipLaunchKernelGGL(k_reduce_grad, dim3((G.nparams + 63) / 64), dim3(256), 0, s, A.partial, nsplit, G.nparams, grad);
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
  if (!params || !grad || !m1 || !m2) return fail(-3, "null pointer%s", "");
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, grad, m1,
                     m2, (int)n, lr, beta1, beta2, eps, bc1, bc2, clip);
  return hip_status(hipGetLastError());
}

int mfx_flow_adam_cast(int D, int L, const int* masks, float* params, uint16_t* params_bf16, const float* grad,
                       float* m1, float* m2, float lr, float beta1, float beta2, float eps, int step, float clip,
                       void* stream) {
  FlowGeom G;
  int rc = make_geom(D, L, masks, G);
  if (rc) return rc;
  if (step < 1) return fail(-2, "step >= 1 required%s", "");
  if (!params || !params_bf16 || !grad || !m1 || !m2) return fail(-3, "null pointer%s", "");
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam_cast, dim3(64, L), dim3(256), 0, (hipStream_t)stream, G, params, grad, m1, m2, lr, beta1,
                     beta2, eps, bc1, bc2, clip, (unsigned short*)params_bf16);
  return hip_status(hipGetLastError());
}


int64_t mfx_flow_cdf_workspace(int64_t n) {
  if (n < 0) return -1;
  return (int64_t)(((n + CDF_B - 1) / CDF_B + 1) * 8);
}

int mfx_flow_cdf(const double* w, int64_t n, double* cdf, double* workspace, int64_t workspace_bytes, void* stream) {
  if (n <= 0) return fail(-2, "n > 0 required%s", "");
  if (!w || !cdf || !workspace) return fail(-3, "null pointer%s", "");
  const int64_t nb = (n + CDF_B - 1) / CDF_B;
  if (nb > (1 << 30) || workspace_bytes < (nb + 1) * 8) return fail(-2, "workspace smaller than mfx_flow_cdf_workspace%s", "");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cdf_blocks, dim3((unsigned)nb), dim3(CDF_T), 0, s, w, n, cdf, workspace);
  hipLaunchKernelGGL(k_cdf_totals, dim3(1), dim3(64), 0, s, workspace, (int)nb);
  hipLaunchKernelGGL(k_cdf_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cdf, n, (const double*)workspace,
                     (int)nb);
  return hip_status(hipGetLastError());
}

int mfx_flow_cdf_guide(const double* cdf, int64_t n, int* guide, int64_t m, void* stream) {
  if (n <= 0 || m <= 0) return fail(-2, "n > 0 and m > 0 required%s", "");
  if (n > 0x7fffffff) return fail(-2, "n must fit in 31 bits%s", "");
  if (!cdf || !guide) return fail(-3, "null pointer%s", "");
  hipLaunchKernelGGL(k_cdf_guide, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream, cdf, n, guide, m);
  return hip_status(hipGetLastError());
}

int mfx_flow_moment(const float* logq, const double* f, int64_t n, double mix, double* work, void* stream) {
  if (n <= 0) return fail(-2, "n > 0 required%s", "");
  if (!logq || !f || !work) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_moment, dim3(MOM_G), dim3(MOM_T), 0, s, logq, f, n, mix, work);
  hipLaunchKernelGGL(k_moment_sum, dim3(1), dim3(64), 0, s, work);
  return hip_status(hipGetLastError());
}

int mfx_flow_mixture_winv(const float* ld, int64_t n, int inverse, double mix, double* winv, void* stream) {
  if (n < 0) return fail(-2, "n >= 0 required%s", "");
  if (n == 0) return 0;
  if (!ld || !winv) return fail(-3, "null pointer%s", "");
  hipLaunchKernelGGL(k_mixture_winv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, ld, n,
                     inverse ? -1.0 : 1.0, mix, winv);
  return hip_status(hipGetLastError());
}

}  // extern "C"
