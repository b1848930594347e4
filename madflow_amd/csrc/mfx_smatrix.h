// Fused smatrix kernels: one launch evaluates, per event (one event per lane), every
// helicity row, the whole diagram program, the jamp accumulation and the colour sum,
// then writes |M|^2 averaged over helicities/colours — nothing intermediate touches
// HBM.  This replaces both the TensorFlow graph of
//   madgraph_plugin/template_files/matrix_method_python.inc:80-138  (smatrix/matrix)
// and the generated custom op, which launched one kernel per helicity
//   python_package/madflow/custom_op/constants.py:164-198, generation.py:96-105.
#pragma once
#include "mfx_helas.h"

namespace mfx {

struct SmatrixArgs {
  const double* all_ps;  // (nev, nexternal, 4) f64, E px py pz
  int64_t nev;
  double mdl_MT, mdl_WT;
  double gc[3][2];       // frozen couplings GC_10, GC_11, GC_12 (re, im)
  const double* gc_ev;   // optional per-event couplings [3][nev][2] (running alpha_s)
  double sqh;            // value of sqrt(1/2) used by vxxxxx
  int helv[8];           // explicit helicity vector (matrix(hel)), when use_helv
  int use_helv;
  double* out;           // (nev,) f64
  double* jamp_out;      // optional (nev, ncolor, 2) f64, only with use_helv
};

template <class R>
__device__ __forceinline__ Params<R> load_params(const SmatrixArgs& a, int64_t e) {
  Params<R> prm;
  prm.mdl_MT = R(a.mdl_MT);
  prm.mdl_WT = R(a.mdl_WT);
  prm.sqh = R(a.sqh);
  if (a.gc_ev) {
    const double* g = a.gc_ev;
    prm.GC_10 = cx<R>(R(g[2 * e]), R(g[2 * e + 1]));
    prm.GC_11 = cx<R>(R(g[2 * (a.nev + e)]), R(g[2 * (a.nev + e) + 1]));
    prm.GC_12 = cx<R>(R(g[2 * (2 * a.nev + e)]), R(g[2 * (2 * a.nev + e) + 1]));
  } else {
    prm.GC_10 = cx<R>(R(a.gc[0][0]), R(a.gc[0][1]));
    prm.GC_11 = cx<R>(R(a.gc[1][0]), R(a.gc[1][1]));
    prm.GC_12 = cx<R>(R(a.gc[2][0]), R(a.gc[2][1]));
  }
  return prm;
}

// sum_ij J_i cf_ij conj(J_j) / den with a real symmetric cf:
//   sum_i cf_ii |J_i|^2 + 2 sum_{i<j} cf_ij Re(J_i conj(J_j))
template <class T, class R>
__device__ __forceinline__ R colour_sum(const cx<R> (&J)[T::NCOLOR]) {
  constexpr int NC = T::NCOLOR;
  R res = R(0);
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    R zr = T::cf(R(), i, i) * J[i].re;
    R zi = T::cf(R(), i, i) * J[i].im;
#pragma unroll
    for (int j = i + 1; j < NC; ++j) {
      const R c = R(2) * T::cf(R(), i, j);
      zr += c * J[j].re;
      zi += c * J[j].im;
    }
    res += zr * J[i].re + zi * J[i].im;
  }
  return res * R(1.0 / T::CF_DEN);
}

// Per-event invariants inv[k] = 2 k_i.k_j (i < j) of the outgoing momenta k_l = sign_l p_l,
// in FP64 from the FP64 inputs: propagator denominators are sums of these, which keeps
// P^2 - M^2 accurate in the FP32 build where it would cancel if formed from components.
template <class T, class R>
__device__ __forceinline__ void load_invariants(const double* __restrict__ all_ps, int64_t e, R (&inv)[T::NINV]) {
  const double* src = all_ps + e * (T::NEXTERNAL * 4);
  double k[T::NEXTERNAL][4];
#pragma unroll
  for (int l = 0; l < T::NEXTERNAL; ++l)
#pragma unroll
    for (int c = 0; c < 4; ++c) k[l][c] = T::leg_sign(l) * src[4 * l + c];
  int idx = 0;
#pragma unroll
  for (int i = 0; i < T::NEXTERNAL; ++i)
#pragma unroll
    for (int j = i + 1; j < T::NEXTERNAL; ++j) {
      inv[idx++] = R(2.0 * (k[i][0] * k[j][0] - k[i][1] * k[j][1] - k[i][2] * k[j][2] - k[i][3] * k[j][3]));
    }
}

template <class T, class R>
__device__ __forceinline__ void load_momenta(const double* __restrict__ all_ps, int64_t e, R (&p)[T::NEXTERNAL][4]) {
  const double* src = all_ps + e * (T::NEXTERNAL * 4);
#pragma unroll
  for (int l = 0; l < T::NEXTERNAL; ++l) {
    const double2 a = *reinterpret_cast<const double2*>(src + 4 * l);
    const double2 b = *reinterpret_cast<const double2*>(src + 4 * l + 2);
    p[l][0] = R(a.x);
    p[l][1] = R(a.y);
    p[l][2] = R(b.x);
    p[l][3] = R(b.y);
  }
}

// The momentum-only parts of the program (slot momenta, propagator denominators)
// are the same for every helicity row; if the compiler hoists them out of the
// helicity loop they all stay live across it and spill.  An empty asm that "may
// modify" the momenta keeps them inside the loop at no instruction cost.
template <class R> __device__ __forceinline__ void opaque(R& x) { asm volatile("" : "+v"(x)); }

// The colour sum runs on the matrix cores for FP32 and 24 or more colour flows.
template <class T, class R> struct MfmaColour {
  static constexpr bool value = StableForms<R>::value && T::NCOLOR >= 24 && T::NCOLOR % 4 == 0;
};

typedef float mfx_cf32x4 __attribute__((ext_vector_type(4)));

// LDS ordering between the lanes of one wave (no block barrier needed).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Re(J^dagger CF J) / CF_DEN of one helicity row for the 64 events of a wave (one event
// per lane), as a batched dense contraction on the matrix cores: the wave stages its
// jamps in LDS (colour-major), Y = J CF runs as (16 events x NCOLOR) x (NCOLOR x 16)
// tiles of v_mfma_f32_16x16x4f32 for Re J and Im J (lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[4(l/16)+r][l%16]), each D tile is dotted with J row by row,
// and the 16 column partial sums of each row are reduced across the quarter-wave.
// stage: 2*NCOLOR*64 floats of this wave; cfl: the integer colour matrix [NCOLOR][NCP];
// red: 64 floats of this wave.
template <class T>
__device__ __forceinline__ float colour_sum_mfma_staged(const float* stage, const float* cfl, float* red);

template <class T>
__device__ __forceinline__ float colour_sum_mfma(const cx<float> (&J)[T::NCOLOR], float* stage, const float* cfl,
                                                 float* red) {
  constexpr int NC = T::NCOLOR;
  const int lane = threadIdx.x & 63;
  wave_lds_sync();                       // the previous row's reads are done
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    stage[(2 * c) * 64 + lane] = J[c].re;
    stage[(2 * c + 1) * 64 + lane] = J[c].im;
  }
  return colour_sum_mfma_staged<T>(stage, cfl, red);
}

// Jamps accumulated in place in a wave's LDS stage (colour-major, one column per lane):
// jamp[c] += amp reads and writes this lane's two words of colour c.
struct LdsCx {
  float* re;
  float* im;
  __device__ __forceinline__ void operator=(const cx<float>& a) { *re = a.re; *im = a.im; }
  __device__ __forceinline__ void operator+=(const cx<float>& a) { *re += a.re; *im += a.im; }
  __device__ __forceinline__ void operator-=(const cx<float>& a) { *re -= a.re; *im -= a.im; }
};
struct LdsJamps {
  float* base;   // stage + lane
  __device__ __forceinline__ LdsCx operator[](int c) const { return LdsCx{base + 2 * c * 64, base + (2 * c + 1) * 64}; }
};

// The contraction of jamps already staged in LDS (see colour_sum_mfma).
template <class T>
__device__ __forceinline__ float colour_sum_mfma_staged(const float* stage, const float* cfl, float* red) {
  constexpr int NC = T::NCOLOR, NCP = (NC + 15) / 16 * 16, KS = NC / 4;
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  wave_lds_sync();
  float part[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int r = 0; r < 4; ++r) part[m][r] = 0.f;
    float are[KS], aim[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      are[k] = stage[(2 * (4 * k + lq)) * 64 + 16 * m + lr];
      aim[k] = stage[(2 * (4 * k + lq) + 1) * 64 + 16 * m + lr];
    }
#pragma unroll
    for (int nt = 0; nt < NCP / 16; ++nt) {
      mfx_cf32x4 dre = {0.f, 0.f, 0.f, 0.f}, dim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        const float b = cfl[(4 * k + lq) * NCP + 16 * nt + lr];
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(are[k], b, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(aim[k], b, dim, 0, 0, 0);
      }
      const int col = 16 * nt + lr;
      if (col < NC) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ev = 16 * m + 4 * lq + r;
          part[m][r] += stage[(2 * col) * 64 + ev] * dre[r] + stage[(2 * col + 1) * 64 + ev] * dim[r];
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) part[m][r] += __shfl_xor(part[m][r], off, 64);
  if (lr == 0) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[16 * m + 4 * lq + r] = part[m][r];
  }
  wave_lds_sync();
  return red[lane] * float(1.0 / T::CF_DEN);
}

template <class T, bool MFMA> struct ColourLds { float stage[1], cf[1], red[1]; };
template <class T> struct ColourLds<T, true> {
  static constexpr int NCP = (T::NCOLOR + 15) / 16 * 16;
  // per wave: the staging of the register jamps, plus one in-place jamp array for each
  // further row of a pass (with four rows, the first row is staged in the second row's
  // array once that row has been contracted)
  static constexpr int NST = T::MULTI_K >= 2 ? 3 : 2;
  float stage[NST][4][2 * T::NCOLOR * 64];
  float cf[T::NCOLOR * NCP];
  float red[4][64];
};

// Hot path: all helicity rows for nev events.  The event loop is block-uniform so that
// every lane of a wave reaches the wave-wide colour contraction; lanes past the end
// evaluate the last event again and write nothing.
template <class T, class R>
__global__ void __launch_bounds__(256) k_smatrix(SmatrixArgs a) {
  constexpr bool MFMA = MfmaColour<T, R>::value;
  __shared__ ColourLds<T, MFMA> lds;
  if constexpr (MFMA) {
    constexpr int NCP = ColourLds<T, true>::NCP;
    for (int i = threadIdx.x; i < T::NCOLOR * NCP; i += blockDim.x) {
      const int r = i / NCP, c = i % NCP;
      lds.cf[i] = c < T::NCOLOR ? T::cf(float(), r, c) : 0.f;
    }
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.nev; base += stride) {
    const int64_t e0 = base + threadIdx.x;
    const bool active = e0 < a.nev;
    const int64_t e = active ? e0 : a.nev - 1;
    R p[T::NEXTERNAL][4];
    R inv[T::NINV];
    load_momenta<T, R>(a.all_ps, e, p);
    if constexpr (StableForms<R>::value) load_invariants<T, R>(a.all_ps, e, inv);
    const Params<R> prm = load_params<R>(a, e);
    R ans = R(0);
    if constexpr (T::MULTI_K >= 1 && StableForms<R>::value) {
      // every helicity combination is a row: walk them in groups of 2^MULTI_K rows that
      // differ only in the helicities of the legs MULTI_LEG, whose shared currents the
      // multi-row program computes once per group
      constexpr int K = T::MULTI_K;
      static_assert(T::NCOMB == (1 << T::NEXTERNAL), "groups of rows need every combination");
      static_assert(K <= 2, "at most four rows per pass");
      constexpr int XA = K == 2 && T::MULTI_LEG[K - 1] < T::MULTI_LEG[0] ? T::MULTI_LEG[K - 1] : T::MULTI_LEG[0];
      constexpr int XB = K == 2 && T::MULTI_LEG[K - 1] < T::MULTI_LEG[0] ? T::MULTI_LEG[0] : T::MULTI_LEG[K - 1];
#pragma unroll 1
      for (int q = 0; q < (T::NCOMB >> K); ++q) {
#pragma unroll
        for (int l = 0; l < T::NEXTERNAL; ++l)
#pragma unroll
          for (int k = 0; k < 4; ++k) opaque(p[l][k]);
#pragma unroll
        for (int k = 0; k < T::NINV; ++k) opaque(inv[k]);
        // zero bits inserted at the paired legs' positions, the lower position first
        unsigned hb = ((unsigned)q & ((1u << XA) - 1u)) | (((unsigned)q >> XA) << (XA + 1));
        if constexpr (K == 2) hb = (hb & ((1u << XB) - 1u)) | ((hb >> XB) << (XB + 1));
        // groups of rows that vanish identically (the paired legs never decide that)
        if (!T::live(hb)) continue;
        if constexpr (K == 2) {
          static_assert(MFMA, "four rows per pass keep three rows' jamps in the colour-sum LDS");
          // rows 1..3 accumulate in place in the wave's LDS arrays; row 0 stays in
          // registers and is staged in row 1's array after row 1 has been contracted
          const int lane = threadIdx.x & 63;
          cx<R> jamp[T::NCOLOR];
          LdsJamps j1{lds.stage[0][wave] + lane}, j2{lds.stage[1][wave] + lane}, j3{lds.stage[2][wave] + lane};
          wave_lds_sync();                 // the previous group's contractions are done
          T::template amplitudes4<R>(p, inv, hb, prm, jamp, j1, j2, j3);
          ans += colour_sum_mfma_staged<T>(lds.stage[0][wave], lds.cf, lds.red[wave]);
          ans += colour_sum_mfma<T>(jamp, lds.stage[0][wave], lds.cf, lds.red[wave]);
          ans += colour_sum_mfma_staged<T>(lds.stage[1][wave], lds.cf, lds.red[wave]);
          ans += colour_sum_mfma_staged<T>(lds.stage[2][wave], lds.cf, lds.red[wave]);
        } else if constexpr (MFMA) {
          // the second row's jamps accumulate in place in the wave's second LDS stage,
          // which frees their registers for the currents the two rows share
          const int lane = threadIdx.x & 63;
          cx<R> jamp[T::NCOLOR];
          LdsJamps jb{lds.stage[1][wave] + lane};
          wave_lds_sync();                 // the previous pair's contractions are done
          T::template amplitudes2<R>(p, inv, hb, prm, jamp, jb);
          ans += colour_sum_mfma<T>(jamp, lds.stage[0][wave], lds.cf, lds.red[wave]);
          ans += colour_sum_mfma_staged<T>(lds.stage[1][wave], lds.cf, lds.red[wave]);
        } else {
          cx<R> jamp[T::NCOLOR], jamp2[T::NCOLOR];
          T::template amplitudes2<R>(p, inv, hb, prm, jamp, jamp2);
          ans += colour_sum<T, R>(jamp);
          ans += colour_sum<T, R>(jamp2);
        }
      }
    } else {
#pragma unroll 1
      for (int h = 0; h < T::NCOMB; ++h) {
#pragma unroll
        for (int l = 0; l < T::NEXTERNAL; ++l)
#pragma unroll
          for (int k = 0; k < 4; ++k) opaque(p[l][k]);
        if constexpr (StableForms<R>::value) {
#pragma unroll
          for (int k = 0; k < T::NINV; ++k) opaque(inv[k]);
        }
        if (!T::live(T::hel_bits(h))) continue;   // a row that vanishes identically
        cx<R> jamp[T::NCOLOR];
        T::template amplitudes<R>(p, inv, T::hel_bits(h), nullptr, prm, jamp);
        if constexpr (MFMA) ans += colour_sum_mfma<T>(jamp, lds.stage[0][wave], lds.cf, lds.red[wave]);
        else ans += colour_sum<T, R>(jamp);
      }
    }
    if (active) a.out[e] = double(ans) / T::DENOMINATOR;
  }
}

// matrix(all_ps, hel, ...): one explicit helicity vector (may hold 4 = BRST check for a
// gluon); optionally also returns the jamps.
template <class T, class R>
__global__ void __launch_bounds__(256) k_matrix_hel(SmatrixArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.nev; e += stride) {
    R p[T::NEXTERNAL][4];
    R inv[T::NINV];
    load_momenta<T, R>(a.all_ps, e, p);
    if constexpr (StableForms<R>::value) load_invariants<T, R>(a.all_ps, e, inv);
    const Params<R> prm = load_params<R>(a, e);
    cx<R> jamp[T::NCOLOR];
    T::template amplitudes<R>(p, inv, 0u, a.helv, prm, jamp);
    a.out[e] = double(colour_sum<T, R>(jamp));
    if (a.jamp_out) {
#pragma unroll
      for (int c = 0; c < T::NCOLOR; ++c) {
        a.jamp_out[(e * T::NCOLOR + c) * 2] = double(jamp[c].re);
        a.jamp_out[(e * T::NCOLOR + c) * 2 + 1] = double(jamp[c].im);
      }
    }
  }
}

struct ProcessEntry {
  const char* name;
  int nexternal, ncomb, ncolor, ndiags, ngraphs;
  double denominator;
  // ws / ws_bytes: caller-owned device workspace (ignored by the fused kernels)
  hipError_t (*launch)(const SmatrixArgs& a, int precision, int per_helicity, void* ws, int64_t ws_bytes,
                       hipStream_t s);
  // workspace bytes per event and helicity row; nullptr: the process needs none
  int64_t (*ws_per_row)(int precision);
};

template <class T>
hipError_t launch_process(const SmatrixArgs& a, int precision, int per_helicity, void*, int64_t, hipStream_t s) {
  if (a.nev <= 0) return hipSuccess;
  const int block = 256;
  int64_t nblk = (a.nev + block - 1) / block;
  if (nblk > (1 << 30)) nblk = 1 << 30;
  dim3 grid((unsigned)nblk), blk(block);
  if (precision == 32) {
    if (per_helicity) hipLaunchKernelGGL((k_matrix_hel<T, float>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_smatrix<T, float>), grid, blk, 0, s, a);
  } else {
    if (per_helicity) hipLaunchKernelGGL((k_matrix_hel<T, double>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_smatrix<T, double>), grid, blk, 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace mfx

namespace mfx {
// Registry filled by the process translation units at library load time.
int register_process(const ProcessEntry& e);
}  // namespace mfx

#define MFX_INSTANTIATE_PROCESS(NAME)                                                                   \
  namespace mfx {                                                                                      \
  static const int registered_##NAME = register_process(                                               \
      ProcessEntry{NAME::Traits::PROC_NAME, NAME::Traits::NEXTERNAL, NAME::Traits::NCOMB,               \
                   NAME::Traits::NCOLOR, NAME::Traits::NDIAGS, NAME::Traits::NGRAPHS,                   \
                   NAME::Traits::DENOMINATOR, &launch_process<NAME::Traits>, nullptr});                 \
  }
