// Colour-ordered evaluation of g g > t t~ + n gluons, for processes whose colour flows
// (n! jamps) do not fit the register file next to a Feynman-diagram program.
//
// The jamps are the coefficients of the fundamental-string colour basis
// (T^{a_s1} ... T^{a_sn})_{i_t j_tbar}, the basis of the generated modules'
// jamp lines (madgraph_plugin/PyOut_exporter.py:334-375).  For one ordering s of the
// gluons the jamp is the colour-ordered amplitude, computed by the Berends-Giele
// recursion over consecutive ranges of s (no Feynman diagram is enumerated):
//
//   J(i..j) = 1/P^2 [ GC_10 sum_{i<=k<j} N3(J(i..k), J(k+1..j))
//                     + (-i GC_12) sum_{i<=k<l<j} N4(J(i..k), J(k+1..l), J(l+1..j)) ]
//   O(0) = top,  O(j) = sum_{i<j} FFV1_1(O(i), J(i..j-1))
//   jamp_s = - sum_{i<n} FFV1_0(anti-top, O(i), J(i..n-1))
//
// N3 is the VVV1 Lorentz structure of VVV1P0_1 without its propagator; the colour-
// ordered coefficient of the f^{abc} vertex (-i for the ordered word) cancels its i.
// N4 = 2 V3 (V2.V4) - V2 (V3.V4) - V4 (V2.V3) is the combination VVVV4 - VVVV1 of the
// SM four-gluon Lorentz structures that the word order selects.  The -1 is the
// fermion-flow factor of the generated jamps.  Summed over the Feynman diagrams
// consistent with an ordering, these give exactly the jamp of that ordering.
//
// Kernel 1 (k_ordered_jamps): one event per lane, every helicity row, every ordering;
// each finished jamp is stored to a workspace [helicity][colour][event] (coalesced),
// so no lane ever holds more than one ordering's currents.
// Kernel 2 (k_colour_mfma): |M|^2 = sum_h Re(J_h^dagger CF J_h) / (CF_DEN * DENOMINATOR)
// as a batched dense contraction on the matrix cores: per helicity, the integer colour
// matrix (NCOLOR x NCOLOR) times a 16-event tile of Re J and Im J (NCOLOR x 16) with
// v_mfma_{f32,f64}_16x16x4, then the column-wise dot product with J.
#pragma once
#include "mfx_smatrix.h"

namespace mfx {

typedef float mfx_f32x4 __attribute__((ext_vector_type(4)));
typedef double mfx_f64x4 __attribute__((ext_vector_type(4)));

// Lorentz structure of VVV1P0_1 without coupling or propagator; V1 slots = V2 + V3.
// With P1 = -(P2 + P3) it reads (V2.V3)(P3 - P2) + V2 (2 V3.P2 + V3.P3) - V3 (V2.P2 + 2 V2.P3).
// Colour-ordered currents of gluons alone are conserved, P.J(i..j) = 0 (external gluons
// on shell and transverse; Berends and Giele, Nucl. Phys. B306 (1988) 759), so the
// V2.P2 and V3.P3 terms vanish and are left out.  d = V2.V3, the dot product the
// structure contains (the four-gluon sums reuse it).
template <class R>
__device__ __forceinline__ void vvv_numerator(const W<R>& V2, const W<R>& V3, W<R>& N, cx<R>& d) {
  R P2[4], P3[4];
  mom_in(V2, P2);
  mom_in(V3, P3);
  N[0] = V2[0] + V3[0];
  N[1] = V2[1] + V3[1];
  const cx<R> TMP2 = mdotp(V3, P2), TMP4 = mdotp(V2, P3);
  const cx<R> TMP5 = mdot(V3, V2);
  const cx<R> c2 = TMP2 + TMP2, c3 = -(TMP4 + TMP4);
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) N[2 + mu] = TMP5 * (P3[mu] - P2[mu]) + V2[2 + mu] * c2 + V3[2 + mu] * c3;
  d = TMP5;
}

// N[2..5] += c V
template <class R>
__device__ __forceinline__ void add_scaled(W<R>& N, const W<R>& V, cx<R> c) {
#pragma unroll
  for (int mu = 2; mu < 6; ++mu) N[mu] += V[mu] * c;
}

template <class R>
__device__ __forceinline__ void add_components(W<R>& acc, const W<R>& t) {
#pragma unroll
  for (int k = 2; k < 6; ++k) acc[k] += t[k];
}

// Colour-ordered currents of one word.  J[i][j] (i <= j) is the gluon current of the
// word positions i..j (J[i][i] = the gluon itself), O[j] the top line carrying the top and
// the first j gluons.  FP32 forms only: P^2 of the gluons at word positions i..j (summed
// where the range's propagator is applied) and rt[j] = P^2 - m_t^2 of the top with the
// first j gluons are sums of the pair invariants pp[a][b] = 2 k_a.k_b (computed in FP64
// per event and rounded once).  Incoming
// momenta enter negated, so these sums can cancel: rt[j] is taken from the shorter of its
// two equal forms, the top with the first j gluons or the anti-top with the others
// (momentum conservation), and of two forms of equal length from the one with the smaller
// sum of magnitudes (summing the top with the first j gluons throughout lost up to 1e-4 of
// |M|^2 in FP32).  All indices are compile-time constants after unrolling, so every
// current lives in registers.
//
// The four-gluon sum of J(i..j) runs over the splits i <= k < l < j into
// V2 = J(i..k), V3 = J(k+1..l), V4 = J(l+1..j).  Two of its three terms only involve dot
// products of adjacent ranges, and summed over one split index those are
// S(a..b) = sum_{a<=m<b} J(a..m).J(m+1..b), the dot products of the three-gluon sum of
// range a..b (kept in S[a][b]):
//   sum_{k<l} V2 (V3.V4) = sum_{i<=k<j-1} J(i..k) S(k+1..j)
//   sum_{k<l} V4 (V2.V3) = sum_{i<l<j}   J(l+1..j) S(i..l)
// so only the 2 V3 (V2.V4) term needs a dot product per split pair.

// J(i..j) (and S(i..j)) for every range ending at j in [J0, J1), shorter ranges first.
template <int NG, int J0, int J1, class R>
__device__ __forceinline__ void ordered_ranges(W<R> (&J)[NG][NG], cx<R> (&S)[NG][NG], const R (&pp)[NG][NG],
                                               const Params<R>& prm) {
  constexpr bool INV = StableForms<R>::value;
  // -i GC_12 multiplies the four-gluon structures (GC_12 = i gs^2: a real gs^2)
  const cx<R> c4 = mul_mi(prm.GC_12);
#pragma unroll
  for (int j = J0; j < J1; ++j) {
#pragma unroll
    for (int i = j - 1; i >= 0; --i) {
      W<R> N3;
      cx<R> sd;
      vvv_numerator(J[i][i], J[i + 1][j], N3, sd);
#pragma unroll
      for (int k = i + 1; k < j; ++k) {
        W<R> t;
        cx<R> d;
        vvv_numerator(J[i][k], J[k + 1][j], t, d);
        add_components(N3, t);
        sd += d;
      }
      S[i][j] = sd;
      W<R> N4;
      N4[0] = N3[0];
      N4[1] = N3[1];
#pragma unroll
      for (int q = 2; q < 6; ++q) N4[q] = cx<R>(R(0), R(0));
#pragma unroll
      for (int k = i; k < j; ++k)
#pragma unroll
        for (int l = k + 1; l < j; ++l) {
          const cx<R> d24 = mdot(J[i][k], J[l + 1][j]);
          add_scaled(N4, J[k + 1][l], d24 + d24);
        }
#pragma unroll
      for (int k = i; k < j - 1; ++k) add_scaled(N4, J[i][k], -S[k + 1][j]);
#pragma unroll
      for (int l = i + 1; l < j; ++l) add_scaled(N4, J[l + 1][j], -S[i][l]);
      // propagator 1/P^2 of the massless gluon
      R den2;
      if constexpr (INV) {
        // P^2 of the gluons at word positions i..j: the sum of their pair invariants
        den2 = R(0);
#pragma unroll
        for (int a = i; a <= j; ++a)
#pragma unroll
          for (int b = a + 1; b <= j; ++b) den2 += pp[a][b];
      } else {
        R P[4];
        mom_out(N3, P);
        den2 = den_from_slots(P, R(0));
      }
      const R rden = recip(den2);
      const cx<R> k3 = prm.GC_10 * rden, k4 = c4 * rden;
      W<R>& o = J[i][j];
      o[0] = N3[0];
      o[1] = N3[1];
#pragma unroll
      for (int q = 2; q < 6; ++q) o[q] = k3 * N3[q] + k4 * N4[q];
    }
  }
}

// O(j) for j in [O0, O1]: O(j) = sum_{i<j} FFV1_1(O(i), J(i..j-1)).  Every term has the
// off-shell momentum of O(j) (top + gluons 0..j-1), so the gamma.J O columns are summed
// and the top propagator is applied once per j.
template <int NG, int O0, int O1, class R>
__device__ __forceinline__ void ordered_top_line(W<R> (&O)[NG], const W<R> (&J)[NG][NG], const R (&rt)[NG],
                                                 const Params<R>& prm) {
  constexpr bool INV = StableForms<R>::value;
#pragma unroll
  for (int j = O0; j <= O1; ++j) {
    VslashF<R> x = vslash_f(O[0], J[0][j - 1]);
#pragma unroll
    for (int i = 1; i < j; ++i) add_col(x, vslash_f(O[i], J[i][j - 1]));
    O[j][0] = O[0][0] + J[0][j - 1][0];
    O[j][1] = O[0][1] + J[0][j - 1][1];
    R P[4];
    mom_out(O[j], P);
    R den2;
    if constexpr (INV) {
      den2 = rt[j];
    } else {
      den2 = den_from_slots(P, prm.mdl_MT);
    }
    ffv1_1_col(x, prm.GC_11, prm.mdl_MT, prm.mdl_WT, den2, P, O[j]);
  }
}

// The closing vertex: jamp = -sum_{i<NG} FFV1_0(anti-top, O(i), J(i..NG-1)).  It is linear
// in the sum of the gamma.J(i..NG-1) O(i) columns, formed once for both anti-top
// helicities (ordered_jamp dots it with each anti-top spinor).
template <int NG, class R>
__device__ __forceinline__ VslashF<R> ordered_close_col(const W<R> (&O)[NG], const W<R> (&J)[NG][NG]) {
  VslashF<R> x = vslash_f(O[0], J[0][NG - 1]);
#pragma unroll
  for (int i = 1; i < NG; ++i) add_col(x, vslash_f(O[i], J[i][NG - 1]));
  return x;
}
template <class R>
__device__ __forceinline__ cx<R> ordered_jamp(const W<R>& atop, const VslashF<R>& x, const Params<R>& prm) {
  return -ffv1_0_col(atop, x, prm.GC_11);
}

// the top-line currents O(1..N-1) of both top helicities (empty when unused)
template <class R, int N, bool ON> struct TopPrefix { W<R> o[2][N]; };
template <class R, int N> struct TopPrefix<R, N, false> {};

struct OrderedArgs {
  SmatrixArgs s;
  int64_t ev0;          // first event of this sub-chunk
  int64_t n;            // events in this sub-chunk
  int64_t ld;           // jbuf row stride in events: n rounded up to a 16-event tile, so that
                        // every tile of a colour row is whole 128-byte lines
  int h0, nhel;         // helicity rows h0 .. h0+nhel-1 (or 1 row for an explicit vector)
  void* jbuf;           // [nhel][NCOLOR][n] complex R
};

// Per-lane staging in LDS (lane-major: slot s of lane t at s * BLOCK + t), used as an
// indexable store: the gluon wavefunctions and the pair invariants are read back in the
// order of each colour ordering (the ordering index is wave-uniform).
template <class T, class R>
struct OrderedStage {
  static constexpr int NG = T::NG;
  static constexpr int NWF = NG * 12;                  // 5 gluons x 6 complex
  // FP32: the pair invariants gluon-gluon and top-gluon (pair_index), then anti-top-gluon
  static constexpr int NPI = StableForms<R>::value ? (NG + 1) * NG / 2 : 0;
  static constexpr int NINV = StableForms<R>::value ? NPI + NG : 0;
  // FP32: the top and anti-top spinors of both helicities (momentum slots once, spinor
  // components per helicity), staged once per event; FP64 recomputes them (its LDS
  // budget is spent)
  static constexpr int FB = NWF + NINV;
  static constexpr int NFS = StableForms<R>::value ? 40 : 0;
  static constexpr int NSLOT = NWF + NINV + NFS;
};

template <class T, class R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_ordered_jamps(OrderedArgs a) {
  constexpr int NG = T::NG, NE = T::NEXTERNAL, NC = T::NCOLOR;
  using S = OrderedStage<T, R>;
  __shared__ R lds[S::NSLOT * BLOCK];
  const int t = threadIdx.x;
  const int64_t el = (int64_t)blockIdx.x * BLOCK + t;
  if (el >= a.n) return;                 // no block-level synchronisation below
  const int64_t e = a.ev0 + el;
  R p[NE][4];
  load_momenta<T, R>(a.s.all_ps, e, p);
  const Params<R> prm = load_params<R>(a.s, e);
  if constexpr (S::NINV > 0) {
    // 2 k_x.k_y in FP64 for x, y among (gluon slots 0..NG-1, top = slot NG)
    const double* src = a.s.all_ps + e * (NE * 4);
    double k[NG + 1][4];
#pragma unroll
    for (int s = 0; s <= NG; ++s) {
      const int leg = s < NG ? T::gluon_leg(s) : T::TOP_LEG;
      const double sg = T::leg_sign(leg);
#pragma unroll
      for (int c = 0; c < 4; ++c) k[s][c] = sg * src[4 * leg + c];
    }
#pragma unroll
    for (int x = 0; x <= NG; ++x)
#pragma unroll
      for (int y = x + 1; y <= NG; ++y) {
        const double v = 2.0 * (k[x][0] * k[y][0] - k[x][1] * k[y][1] - k[x][2] * k[y][2] - k[x][3] * k[y][3]);
        lds[(S::NWF + T::pair_index(x, y)) * BLOCK + t] = R(v);
      }
    // anti-top with each gluon slot (outgoing anti-top: sign +1)
    const double* ka = src + 4 * T::ATOP_LEG;
#pragma unroll
    for (int x = 0; x < NG; ++x) {
      const double v = 2.0 * (k[x][0] * ka[0] - k[x][1] * ka[1] - k[x][2] * ka[2] - k[x][3] * ka[3]);
      lds[(S::NWF + S::NPI + x) * BLOCK + t] = R(v);
    }
  }
  if constexpr (S::NFS > 0) {
    const int* hv = a.s.use_helv ? a.s.helv : nullptr;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      W<R> top, atop;
      T::template fermions<R>(p, ((unsigned)h << T::TOP_LEG) | ((unsigned)h << T::ATOP_LEG), hv, prm, top, atop);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int ot = q < 2 ? 2 * q : 8 + 8 * h + 2 * (q - 2);       // top
        const int oa = q < 2 ? 4 + 2 * q : 24 + 8 * h + 2 * (q - 2);  // anti-top
        if (q >= 2 || h == 0) {
          lds[(S::FB + ot) * BLOCK + t] = top[q].re;
          lds[(S::FB + ot + 1) * BLOCK + t] = top[q].im;
          lds[(S::FB + oa) * BLOCK + t] = atop[q].re;
          lds[(S::FB + oa + 1) * BLOCK + t] = atop[q].im;
        }
      }
    }
  }
  // spinor of the top (anti = 0) or anti-top (anti = 1) at helicity bit h, from LDS
  auto staged_fermion = [&](int anti, int h, W<R>& w) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int o = q < 2 ? 4 * anti + 2 * q : 8 + 16 * anti + 8 * h + 2 * (q - 2);
      w[q].re = lds[(S::FB + o) * BLOCK + t];
      w[q].im = lds[(S::FB + o + 1) * BLOCK + t];
    }
  };
  cx<R>* jb = reinterpret_cast<cx<R>*>(a.jbuf);
  // Helicity rows are enumerated here as r = 4 g + 2 ht + ha: g = the gluon helicities
  // (bit s = gluon slot s), ht / ha = the top / anti-top helicity.  The gluon currents of
  // an ordering depend on g alone, so they are computed once for the four rows of g; the
  // top line is linear in the top spinor and is computed once per ht; the closing
  // vertex is linear in the anti-top spinor.  Rows r0 .. r1-1 are this launch's, each
  // stored at slot r - r0.  With an explicit helicity vector there is one row (r = 0),
  // and every helicity comes from the vector.
  static_assert(T::NCOMB == (4 << NG), "every helicity combination is a row");
  const int r0 = a.h0, r1 = a.h0 + a.nhel;
  constexpr int PRE = NG - 2;
  W<R> J[NG][NG], O[NG];
  cx<R> SD[NG][NG];   // adjacent-range dot sums S(a..b) of the four-gluon terms
  // FP32, five or more gluons: the top-line currents O(1..PRE) of both top helicities,
  // kept from the first ordering of a pair for the second.  Not for FP64 (no registers
  // to spare) nor for four gluons (the kernel fits two waves per SIMD without them).
  constexpr bool SHARE_TOP = S::NFS > 0 && NG >= 5;
  TopPrefix<R, PRE + 1, SHARE_TOP> OP;
  // FP32 forms: the word's pair invariants (gluon-gluon pp, top-gluon tp, anti-top-gluon
  // ap) and the top-line denominators rt (above); pp of the prefix pairs carries over to
  // the second ordering of a pair
  R pp[NG][NG], tp[NG], ap[NG], rt[NG];
#pragma unroll 1
  for (int g = r0 >> 2; g <= (r1 - 1) >> 2; ++g) {
    unsigned hbg = 0u;
#pragma unroll
    for (int s = 0; s < NG; ++s) hbg |= ((unsigned)(g >> s) & 1u) << T::gluon_leg(s);
    const int* helv = a.s.use_helv ? a.s.helv : nullptr;
#pragma unroll
    for (int s = 0; s < NG; ++s) {
      W<R> gw;
      T::template gluon<R>(p, s, hbg, helv, prm, gw);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        lds[(s * 12 + 2 * q) * BLOCK + t] = gw[q].re;
        lds[(s * 12 + 2 * q + 1) * BLOCK + t] = gw[q].im;
      }
    }
    // Orderings come in lexicographic order, so orderings 2m and 2m+1 share their first
    // NG-2 gluons: the currents inside that prefix (J(i..j), j < NG-2) are computed once
    // per pair and carried to the second ordering.
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      int perm[NG];
#pragma unroll
      for (int k = 0; k < NG; ++k) perm[k] = T::perm(c, k);
      const bool fresh = (c & 1) == 0;
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        if (k < PRE && !fresh) continue;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          J[k][k][q].re = lds[(perm[k] * 12 + 2 * q) * BLOCK + t];
          J[k][k][q].im = lds[(perm[k] * 12 + 2 * q + 1) * BLOCK + t];
        }
      }
      if constexpr (S::NINV > 0) {
        // the word's pair invariants (the prefix pairs of a pair's second ordering are the
        // first's), and rt[j] for the top with the first j gluons from the shorter of its
        // two forms
#pragma unroll
        for (int x = 0; x < NG; ++x) {
          tp[x] = lds[(S::NWF + T::pair_index(perm[x], NG)) * BLOCK + t];
          ap[x] = lds[(S::NWF + S::NPI + perm[x]) * BLOCK + t];
#pragma unroll
          for (int y = x + 1; y < NG; ++y) {
            if (y < PRE && !fresh) continue;
            pp[x][y] = lds[(S::NWF + T::pair_index(perm[x], perm[y])) * BLOCK + t];
          }
        }
#pragma unroll
        for (int j = 1; j < NG; ++j) {
          // direct: the top and gluons 0..j-1; complementary: the anti-top and gluons j..NG-1
          const int nd = j + j * (j - 1) / 2, nc = (NG - j) + (NG - j) * (NG - j - 1) / 2;
          R sd = R(0), md = R(0), sc = R(0), mc = R(0);
#pragma unroll
          for (int a = 0; a < j; ++a) {
            sd += tp[a];
            md += fabs(tp[a]);
#pragma unroll
            for (int b = a + 1; b < j; ++b) {
              sd += pp[a][b];
              md += fabs(pp[a][b]);
            }
          }
#pragma unroll
          for (int a = j; a < NG; ++a) {
            sc += ap[a];
            mc += fabs(ap[a]);
#pragma unroll
            for (int b = a + 1; b < NG; ++b) {
              sc += pp[a][b];
              mc += fabs(pp[a][b]);
            }
          }
          rt[j] = nd < nc ? sd : (nc < nd ? sc : (md <= mc ? sd : sc));
        }
      }
      if (fresh) ordered_ranges<NG, 1, PRE, R>(J, SD, pp, prm);
      ordered_ranges<NG, PRE, NG, R>(J, SD, pp, prm);
      if constexpr (SHARE_TOP) {
        // unrolled: OP is indexed by ht
#pragma unroll
        for (int ht = 0; ht < 2; ++ht) {
          const int rb = 4 * g + 2 * ht;
          if (rb + 1 < r0 || rb >= r1) continue;
          W<R> top, atop;
          if constexpr (S::NFS > 0) staged_fermion(0, ht, top);
          else T::template fermions<R>(p, hbg | ((unsigned)ht << T::TOP_LEG), helv, prm, top, atop);
          O[0] = top;
          // O(1..PRE) only involve the prefix gluons: computed once per pair of orderings
          if (fresh) {
            ordered_top_line<NG, 1, PRE, R>(O, J, rt, prm);
#pragma unroll
            for (int j = 1; j <= PRE; ++j) OP.o[ht][j] = O[j];
          } else {
#pragma unroll
            for (int j = 1; j <= PRE; ++j) O[j] = OP.o[ht][j];
          }
          ordered_top_line<NG, PRE + 1, NG - 1, R>(O, J, rt, prm);
          const VslashF<R> x = ordered_close_col<NG, R>(O, J);
#pragma unroll 1
          for (int ha = 0; ha < 2; ++ha) {
            const int r = rb + ha;
            if (r < r0 || r >= r1) continue;
            if constexpr (S::NFS > 0) staged_fermion(1, ha, atop);
            else if (ha) T::template fermions<R>(p, hbg | ((unsigned)ht << T::TOP_LEG) | (1u << T::ATOP_LEG), helv,
                                                 prm, top, atop);
            jb[((int64_t)(r - r0) * NC + c) * a.ld + el] = ordered_jamp(atop, x, prm);
          }
        }
      } else {
#pragma unroll 1
        for (int ht = 0; ht < 2; ++ht) {
          const int rb = 4 * g + 2 * ht;
          if (rb + 1 < r0 || rb >= r1) continue;
          W<R> top, atop;
          if constexpr (S::NFS > 0) staged_fermion(0, ht, top);
          else T::template fermions<R>(p, hbg | ((unsigned)ht << T::TOP_LEG), helv, prm, top, atop);
          O[0] = top;
          ordered_top_line<NG, 1, NG - 1, R>(O, J, rt, prm);
          const VslashF<R> x = ordered_close_col<NG, R>(O, J);
#pragma unroll 1
          for (int ha = 0; ha < 2; ++ha) {
            const int r = rb + ha;
            if (r < r0 || r >= r1) continue;
            if constexpr (S::NFS > 0) staged_fermion(1, ha, atop);
            else if (ha) T::template fermions<R>(p, hbg | ((unsigned)ht << T::TOP_LEG) | (1u << T::ATOP_LEG), helv,
                                                 prm, top, atop);
            jb[((int64_t)(r - r0) * NC + c) * a.ld + el] = ordered_jamp(atop, x, prm);
          }
        }
      }
    }
  }
}

// MFMA 16x16x4 on one wave: D += A(16x4) * B(4x16); lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[drow(l/16, r)][l%16], r = 0..3, where the row is
// 4*(l/16) + r for f32 and (l/16) + 4*r for f64 (measured on gfx950).
__device__ __forceinline__ mfx_f32x4 mfma16x16x4(float a, float b, mfx_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mfx_f64x4 mfma16x16x4(double a, double b, mfx_f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <class R> __device__ __forceinline__ int drow(int lq, int r);
template <> __device__ __forceinline__ int drow<float>(int lq, int r) { return 4 * lq + r; }
template <> __device__ __forceinline__ int drow<double>(int lq, int r) { return lq + 4 * r; }
template <class R> struct Vec4;
template <> struct Vec4<float> { typedef mfx_f32x4 type; };
template <> struct Vec4<double> { typedef mfx_f64x4 type; };

// |M|^2 of the events of one sub-chunk from their jamps: one wave per 16-event tile,
// four tiles per block.  Per helicity and tile the matrix cores form D = CF J^T
// (NCP x 16: colour rows, event columns) from the integer colour matrix (A operand, in
// LDS) and the tile's jamps (B operand: lane (i, q) holds J[event i][colour 4 s + q] for
// every k-step s), then each lane dots its D entries with jamps it already holds.  For
// that, the colour of A row drow(q, r) of row tile mt is 4 (4 mt + r) + q: the output the
// lane holds in register r is then the row of the colour of its k-step 4 mt + r, and the
// jamps are read once.  The four lane quarters of each event are summed at the end.
template <class T, class R>
__global__ void __launch_bounds__(256) k_colour_mfma(const void* jbuf, int64_t n, int64_t ld, int nhel,
                                                     double scale, int accumulate, double* out) {
  constexpr int NC = T::NCOLOR;
  constexpr int NCP = (NC + 15) / 16 * 16;    // padded colour rows of D
  constexpr int KS = NC / 4;                  // k-steps (NCOLOR is a multiple of 4)
  static_assert(NC % 4 == 0, "NCOLOR must be a multiple of 4");
  typedef typename Vec4<R>::type V4;
  __shared__ float cf[NCP * NC];   // integer numerators, rows permuted as above: exact in float
  for (int i = threadIdx.x; i < NCP * NC; i += 256) {
    const int row = i / NC, k = i % NC;
    const int mt = row / 16, x = row % 16;
    // x = drow(q, r): 4 q + r for f32, q + 4 r for f64
    const int q = drow<R>(1, 0) == 4 ? x / 4 : x % 4, r = drow<R>(1, 0) == 4 ? x % 4 : x / 4;
    const int colour = 4 * (4 * mt + r) + q;
    cf[i] = colour < NC ? T::cf(float(), colour, k) : 0.0f;
  }
  __syncthreads();
  const cx<R>* J = reinterpret_cast<const cx<R>*>(jbuf);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int64_t ntiles = (n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t ea = tile * 16 + lr;          // the event of this lane's B column
    // each helicity's colour terms are summed in R, the helicities in FP64 (an FP32 sum
    // over every helicity and colour would lose digits with the number of rows)
    double acc = 0.0;
    for (int h = 0; h < nhel; ++h) {
      const cx<R>* Jh = J + (int64_t)h * NC * ld;
      R are[KS], aim[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const cx<R> v = ea < n ? Jh[(int64_t)(4 * s + lq) * ld + ea] : cx<R>(R(0), R(0));
        are[s] = v.re;
        aim[s] = v.im;
      }
      R part = R(0);
#pragma unroll
      for (int mt = 0; mt < NCP / 16; ++mt) {
        V4 dre = {R(0), R(0), R(0), R(0)}, dim = {R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const R c = R(cf[(16 * mt + lr) * NC + 4 * s + lq]);
          dre = mfma16x16x4(c, are[s], dre);
          dim = mfma16x16x4(c, aim[s], dim);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * mt + r < KS) part += are[4 * mt + r] * dre[r] + aim[4 * mt + r] * dim[r];
        // one row tile at a time: interleaving the tiles would hold every tile's
        // accumulators at once
        __builtin_amdgcn_sched_barrier(0);
      }
      acc += double(part);
    }
    // sum the four lane quarters that hold the colours of each event
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    if (lq == 0 && ea < n) out[ea] = (accumulate ? out[ea] : 0.0) + acc * scale;
  }
}

// jamps of a single helicity row [NCOLOR][ld] -> (n, NCOLOR, 2) f64
template <class R>
__global__ void k_jamp_transpose(const void* jbuf, int64_t n, int64_t ld, int ncolor, double* jamp_out) {
  const cx<R>* J = reinterpret_cast<const cx<R>*>(jbuf);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * ncolor) return;
  const int64_t e = i / ncolor;
  const int c = (int)(i % ncolor);
  const cx<R> v = J[(int64_t)c * ld + e];
  jamp_out[2 * i] = double(v.re);
  jamp_out[2 * i + 1] = double(v.im);
}

// bytes of jamp workspace per event and helicity row
template <class T>
int64_t ordered_ws_per_row(int precision) {
  return (int64_t)T::NCOLOR * 2 * (precision == 32 ? 4 : 8);
}

// Sub-chunks of events and groups of helicity rows sized to the caller's workspace:
// the largest helicity group that still leaves a sub-chunk of min(nev, 65536) events
// (enough waves for every SIMD), then as many events as the workspace holds.
template <class T>
hipError_t launch_ordered(const SmatrixArgs& s, int precision, int per_helicity, void* ws, int64_t ws_bytes,
                          hipStream_t st) {
  if (s.nev <= 0) return hipSuccess;
  const int nrows = per_helicity ? 1 : T::NCOMB;
  const int64_t row = ordered_ws_per_row<T>(precision);
  const int64_t want = s.nev < 65536 ? s.nev : 65536;
  int hg = nrows;
  while (hg > 1 && ws_bytes / (row * hg) < want) hg /= 2;
  // whole 16-event tiles per pass: the row stride (n rounded up to 16) then fits too
  int64_t nsub = ws ? ws_bytes / (row * hg) / 16 * 16 : 0;
  if (nsub < 16) return hipErrorInvalidValue;   // the entry point checks the size first
  // Passes of a whole number of waves per SIMD (4 SIMDs per CU, one event per lane) when a
  // pass holds few: at one wave per SIMD a pass of 1.07 waves per SIMD takes two wave-times,
  // with most SIMDs idle in the second.  From 8 waves per SIMD on, the partial last round
  // costs less than the extra pass that rounding down would add.
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const int64_t fill = (int64_t)cus * 4 * 64;
  if (nsub > fill && nsub < 8 * fill) nsub = nsub / fill * fill;
  const double scale = 1.0 / (T::CF_DEN * (per_helicity ? 1.0 : T::DENOMINATOR));
  for (int64_t off = 0; off < s.nev; off += nsub) {
    const int64_t n = s.nev - off < nsub ? s.nev - off : nsub;
    const int64_t tiles = (n + 15) / 16;
    int64_t cblk = (tiles + 3) / 4;
    if (cblk > 65536) cblk = 65536;
    for (int h0 = 0; h0 < nrows; h0 += hg) {
      OrderedArgs a;
      a.s = s;
      a.ev0 = off;
      a.n = n;
      a.ld = (n + 15) / 16 * 16;
      a.h0 = h0;
      a.nhel = hg;
      a.jbuf = ws;
      const int acc = h0 > 0;
      if (precision == 32) {
        hipLaunchKernelGGL((k_ordered_jamps<T, float, 64>), dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, a);
        hipLaunchKernelGGL((k_colour_mfma<T, float>), dim3((unsigned)cblk), dim3(256), 0, st, (const void*)ws, n, a.ld,
                           hg, scale, acc, s.out + off);
      } else {
        hipLaunchKernelGGL((k_ordered_jamps<T, double, 64>), dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, a);
        hipLaunchKernelGGL((k_colour_mfma<T, double>), dim3((unsigned)cblk), dim3(256), 0, st, (const void*)ws, n,
                           a.ld, hg, scale, acc, s.out + off);
      }
    }
    if (per_helicity && s.jamp_out) {
      const unsigned nb = (unsigned)((n * T::NCOLOR + 255) / 256);
      const int64_t ld = (n + 15) / 16 * 16;
      if (precision == 32)
        hipLaunchKernelGGL((k_jamp_transpose<float>), dim3(nb), dim3(256), 0, st, (const void*)ws, n, ld, T::NCOLOR,
                           s.jamp_out + off * T::NCOLOR * 2);
      else
        hipLaunchKernelGGL((k_jamp_transpose<double>), dim3(nb), dim3(256), 0, st, (const void*)ws, n, ld, T::NCOLOR,
                           s.jamp_out + off * T::NCOLOR * 2);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

}  // namespace mfx

#define MFX_INSTANTIATE_ORDERED(NAME)                                                                   \
  namespace mfx {                                                                                      \
  static const int registered_##NAME = register_process(ProcessEntry{                                  \
      NAME::Traits::PROC_NAME, NAME::Traits::NEXTERNAL, NAME::Traits::NCOMB, NAME::Traits::NCOLOR,      \
      NAME::Traits::NDIAGS, NAME::Traits::NGRAPHS, NAME::Traits::DENOMINATOR, &launch_ordered<NAME::Traits>, \
      &ordered_ws_per_row<NAME::Traits>});                                                           \
  }
