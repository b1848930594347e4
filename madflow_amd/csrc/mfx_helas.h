// HELAS external wavefunctions and ALOHA vertex/propagator routines as gfx950
// device functions (one event per lane, R = float or double).
//
// Conventions follow madflow's TensorFlow routines:
//   ixxxxx / oxxxxx / vxxxxx  python_package/madflow/wavefunctions_flow.py:54-154
//   FFV1_0 / FFV1_1 / FFV1_2 / VVV1P0_1  python_package/madflow/tests/mockup_debug_me.py:34-397
//   FFV1P0_3, VVV1_0, VVVV{1,3,4}_0, VVVV{1,3,4}P0_1: SM UFO Lorentz structures written
//   with the same ALOHA conventions (madgraph_plugin/PyOut_create_aloha.py:124-328):
//   momenta come from slots [0],[1] (+ for inputs, - for the off-shell leg),
//   off-shell currents are COUP/(P^2 - M(M - iW)) times the Lorentz structure
//   (times i for fermions, -i for vectors), amplitudes are COUP * (-i) * L.
//
// The vertex expressions are regrouped around the light-cone combinations
// V0 +- V3, V1 +- i V2 and P0 +- P3, P1 +- i P2 so that each spinor component costs a
// few complex FMAs; branch choices that depend on the event (p at rest, p along -z,
// pT = 0) are per-lane selects, choices that depend on mass/helicity are uniform.
#pragma once
#include "mfx_core.h"

namespace mfx {

template <class R> __device__ __forceinline__ R sel(bool c, R a, R b) { return c ? a : b; }
template <class R> __device__ __forceinline__ R tfsign(R y) { return y > R(0) ? R(1) : (y < R(0) ? R(-1) : R(0)); }

// |p| + p_z of a spinor (pp = |p| for massive, p0 for massless fermions).  FP64: the
// reference's expression max(pp + pz, 0).  FP32: for pz < 0 the equivalent
// pT^2 / (pp - pz), which has no cancellation when the particle moves close to -z
// (pp + pz -> 0 loses all FP32 digits there).
template <class R> __device__ __forceinline__ R plus_component(const R (&p)[4], R pp) {
  if constexpr (StableForms<R>::value) {
    return p[3] >= R(0) ? pp + p[3] : (p[1] * p[1] + p[2] * p[2]) / (pp - p[3]);
  } else {
    return fmax(pp + p[3], R(0));
  }
}

// ------------------------------------------------------------------------ ixxxxx
template <class R>
__device__ __forceinline__ void ixxxxx(const R (&p)[4], R fmass, R nhel, R nsf, W<R>& fi) {
  fi[0] = cx<R>(-p[0] * nsf, -p[3] * nsf);
  fi[1] = cx<R>(-p[1] * nsf, -p[2] * nsf);
  const R nh = nhel * nsf;
  if (fmass != R(0)) {
    const R pp = fmin(p[0], sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]));
    const bool hpos = nh > R(0);  // ip = 1, im = 0 ; else ip = 0, im = 1
    // p at rest (_ox_massive_pp_zero with the projectors exchanged)
    const R sqm0 = sqrt(fabs(fmass));
    const R sqm1 = sqm0 * tfsign(fmass);
    const R ipr = hpos ? R(1) : R(0), imr = hpos ? R(0) : R(1);
    const R sq_ip = hpos ? sqm1 : sqm0, sq_im = hpos ? sqm0 : sqm1;
    const R z0 = ipr * sq_ip, z1 = imr * nsf * sq_ip, z2 = ipr * nsf * sq_im, z3 = imr * sq_im;
    // general case (_ix_massive_pp_nonzero)
    const R sf0 = (R(1) + nsf + (R(1) - nsf) * nh) * R(0.5);
    const R sf1 = (R(1) + nsf - (R(1) - nsf) * nh) * R(0.5);
    const R om0 = sqrt(p[0] + pp);
    const R om1 = fmass / om0;
    const R sfomeg0 = sf0 * (hpos ? om1 : om0);
    const R sfomeg1 = sf1 * (hpos ? om0 : om1);
    const R pp3 = plus_component(p, pp);
    const R d = sqrt(R(2) * pp * pp3);
    const bool z = pp3 == R(0);
    const cx<R> chi1(sel(z, -nh, nh * p[1] / d), sel(z, R(0), p[2] / d));
    const cx<R> chi0(sqrt(pp3 * R(0.5) / pp), R(0));
    const cx<R> cim = hpos ? chi0 : chi1, cip = hpos ? chi1 : chi0;
    const bool rest = pp == R(0);
    fi[2] = rest ? cx<R>(z0, R(0)) : cim * sfomeg0;
    fi[3] = rest ? cx<R>(z1, R(0)) : cip * sfomeg0;
    fi[4] = rest ? cx<R>(z2, R(0)) : cim * sfomeg1;
    fi[5] = rest ? cx<R>(z3, R(0)) : cip * sfomeg1;
  } else {
    const R sqp0p3 = sqrt(plus_component(p, p[0])) * nsf;
    const bool z = sqp0p3 == R(0);
    const cx<R> chi1(sel(z, -nhel * sqrt(R(2) * p[0]), nh * p[1] / sqp0p3), sel(z, R(0), p[2] / sqp0p3));
    const cx<R> chi0(sqp0p3, R(0));
    const cx<R> zero(R(0), R(0));
    if (nh == R(1)) {
      fi[2] = zero; fi[3] = zero; fi[4] = chi0; fi[5] = chi1;
    } else {
      fi[2] = chi1; fi[3] = chi0; fi[4] = zero; fi[5] = zero;
    }
  }
}

// ------------------------------------------------------------------------ oxxxxx
template <class R>
__device__ __forceinline__ void oxxxxx(const R (&p)[4], R fmass, R nhel, R nsf, W<R>& fo) {
  fo[0] = cx<R>(p[0] * nsf, p[3] * nsf);
  fo[1] = cx<R>(p[1] * nsf, p[2] * nsf);
  const R nh = nhel * nsf;
  if (fmass != R(0)) {
    const R pp = fmin(p[0], sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]));
    const bool hpos = nh > R(0);
    // p at rest: ip = -((1-nh)//2)*nhel, im = (1+nh)//2*nhel (integers in {-1,0,1})
    const R ipa = hpos ? R(0) : -nhel;
    const R ima = hpos ? nhel : R(0);
    const R sqm0 = sqrt(fabs(fmass));
    const R sqm1 = sqm0 * tfsign(fmass);
    const R s_ima = (ima != R(0)) ? sqm1 : sqm0;
    const R s_ipa = (ipa != R(0)) ? sqm1 : sqm0;
    const R z0 = ima * s_ima, z1 = ipa * nsf * s_ima, z2 = ima * nsf * s_ipa, z3 = ipa * s_ipa;
    // general case (_ox_massive_pp_nonzero)
    const R sf0 = (R(1) + nsf + (R(1) - nsf) * nh) * R(0.5);
    const R sf1 = (R(1) + nsf - (R(1) - nsf) * nh) * R(0.5);
    const R om0 = sqrt(p[0] + pp);
    const R om1 = fmass / om0;
    const R sfomeg0 = sf0 * (hpos ? om1 : om0);
    const R sfomeg1 = sf1 * (hpos ? om0 : om1);
    const R pp3 = plus_component(p, pp);
    const R d = sqrt(R(2) * pp * pp3);
    const bool z = pp3 == R(0);
    const cx<R> chi1(sel(z, -nh, nh * p[1] / d), sel(z, R(0), -p[2] / d));
    const cx<R> chi0(sqrt(pp3 * R(0.5) / pp), R(0));
    const cx<R> cim = hpos ? chi0 : chi1, cip = hpos ? chi1 : chi0;
    const bool rest = pp == R(0);
    fo[2] = rest ? cx<R>(z0, R(0)) : cim * sfomeg1;
    fo[3] = rest ? cx<R>(z1, R(0)) : cip * sfomeg1;
    fo[4] = rest ? cx<R>(z2, R(0)) : cim * sfomeg0;
    fo[5] = rest ? cx<R>(z3, R(0)) : cip * sfomeg0;
  } else {
    const R sqp0p3 = sqrt(plus_component(p, p[0])) * nsf;
    const bool z = sqp0p3 == R(0);
    const cx<R> chi0(sel(z, -nhel * sqrt(R(2) * p[0]), nh * p[1] / sqp0p3), sel(z, R(0), -p[2] / sqp0p3));
    const cx<R> chi1(sqp0p3, R(0));
    const cx<R> zero(R(0), R(0));
    if (nh == R(1)) {
      fo[2] = chi1; fo[3] = chi0; fo[4] = zero; fo[5] = zero;
    } else {
      fo[2] = zero; fo[3] = zero; fo[4] = chi0; fo[5] = chi1;
    }
  }
}

// ------------------------------------------------------------------------ sxxxxx
// Scalar external wavefunction (wavefunctions_flow.py:32-51): the two momentum slots
// and the constant 1.  Only s[0..2] are meaningful.
template <class R>
__device__ __forceinline__ void sxxxxx(const R (&p)[4], R nss, W<R>& s) {
  s[0] = cx<R>(p[0] * nss, p[3] * nss);
  s[1] = cx<R>(p[1] * nss, p[2] * nss);
  s[2] = cx<R>(R(1), R(0));
  s[3] = s[4] = s[5] = cx<R>(R(0), R(0));
}

// ------------------------------------------------------------------------ vxxxxx
// nhel = 4 gives the BRST (gauge) check polarisation eps = p/p0 (p/vmass if massive).
template <class R>
__device__ __forceinline__ void vxxxxx(const R (&p)[4], R vmass, R nhel, R nsv, R SQH, W<R>& eps) {
  eps[0] = cx<R>(p[0] * nsv, p[3] * nsv);
  eps[1] = cx<R>(p[1] * nsv, p[2] * nsv);
  if (nhel == R(4)) {
    const R d = (vmass == R(0)) ? p[0] : vmass;
    for (int k = 0; k < 4; ++k) eps[2 + k] = cx<R>(p[k] / d, R(0));
    return;
  }
  if (vmass != R(0)) {
    const R pt2 = p[1] * p[1] + p[2] * p[2];
    const R pp = fmin(p[0], sqrt(pt2 + p[3] * p[3]));
    const R pt = fmin(pp, sqrt(pt2));
    const R hel0 = R(1) - fabs(nhel);
    const R nsvahl = nsv * fabs(nhel);
    const R emp = p[0] / (vmass * pp);
    const bool rest = pp == R(0);
    const bool ptz = pt == R(0);
    const R pzpt = p[3] / (pp * pt) * SQH * nhel;
    const cx<R> a1(hel0 * p[1] * emp - p[1] * pzpt, -nsvahl * p[2] / pt * SQH);
    const cx<R> a2(hel0 * p[2] * emp - p[2] * pzpt, nsvahl * p[1] / pt * SQH);
    const cx<R> b1(-nhel * SQH, R(0));
    const cx<R> b2(R(0), nsvahl * SQH * tfsign(p[3]));
    const cx<R> g0(hel0 * pp / vmass, R(0));
    const cx<R> g3(hel0 * p[3] * emp + nhel * pt / pp * SQH, R(0));
    eps[2] = rest ? cx<R>(R(1), R(0)) : g0;
    eps[3] = rest ? cx<R>(-nhel * SQH, R(0)) : (ptz ? b1 : a1);
    eps[4] = rest ? cx<R>(R(0), nsvahl * SQH) : (ptz ? b2 : a2);
    eps[5] = rest ? cx<R>(hel0, R(0)) : g3;
  } else {
    const R pp = p[0];
    const R pt = sqrt(p[1] * p[1] + p[2] * p[2]);
    const bool ptz = pt == R(0);
    const R pzpt = p[3] / (pp * pt) * SQH * nhel;
    const R ipt = R(1) / pt;
    eps[2] = cx<R>(R(0), R(0));
    eps[3] = ptz ? cx<R>(-nhel * SQH, R(0)) : cx<R>(-p[1] * pzpt, -nsv * p[2] * ipt * SQH);
    eps[4] = ptz ? cx<R>(R(0), nsv * SQH * tfsign(p[3])) : cx<R>(-p[2] * pzpt, nsv * p[1] * ipt * SQH);
    eps[5] = cx<R>(nhel * pt / pp * SQH, R(0));
  }
}

// ------------------------------------------------------------------ ALOHA helpers
template <class R> __device__ __forceinline__ void mom_in(const W<R>& w, R (&P)[4]) {
  P[0] = w[0].re; P[1] = w[1].re; P[2] = w[1].im; P[3] = w[0].im;
}
template <class R> __device__ __forceinline__ void mom_out(const W<R>& w, R (&P)[4]) {
  P[0] = -w[0].re; P[1] = -w[1].re; P[2] = -w[1].im; P[3] = -w[0].im;
}
// P^2 - M^2 from the slot momentum of the current
template <class R> __device__ __forceinline__ R den_from_slots(const R (&P)[4], R M) {
  return P[0] * P[0] - P[1] * P[1] - P[2] * P[2] - P[3] * P[3] - M * M;
}
// COUP / (P^2 - M^2 + i M W), with P^2 - M^2 given
template <class R> __device__ __forceinline__ cx<R> prop_denom(cx<R> coup, R den2, R M, R Wd) {
  return cdiv(coup, cx<R>(den2, M * Wd));
}
template <class R> __device__ __forceinline__ cx<R> mdot(const W<R>& a, const W<R>& b) {
  return a[2] * b[2] - a[3] * b[3] - a[4] * b[4] - a[5] * b[5];
}
template <class R> __device__ __forceinline__ cx<R> mdotp(const W<R>& a, const R (&P)[4]) {
  return a[2] * P[0] - a[3] * P[1] - a[4] * P[2] - a[5] * P[3];
}
// light-cone combinations of a vector wavefunction: a = V0+V3, b = V0-V3, c = V1+iV2, d = V1-iV2
template <class R> struct LC { cx<R> a, b, c, d; };
template <class R> __device__ __forceinline__ LC<R> lightcone(const W<R>& V) {
  LC<R> l;
  l.a = V[2] + V[5];
  l.b = V[2] - V[5];
  l.c = V[3] + mul_i(V[4]);
  l.d = V[3] - mul_i(V[4]);
  return l;
}

// ------------------------------------------------------------------------ FFV1_0
// The fermion bilinear F1bar gamma^mu F2 in light-cone components (the coefficients of
// V0+V3, V0-V3, V1+iV2, V1-iV2): a program that closes many diagrams on the same fermion
// pair evaluates it once and each amplitude costs four complex products.
template <class R> struct Bilinear { cx<R> a, b, c, d; };
template <class R>
__device__ __forceinline__ Bilinear<R> bilinear(const W<R>& F1, const W<R>& F2) {
  Bilinear<R> j;
  j.a = F1[2] * F2[4] + F1[5] * F2[3];
  j.b = F1[3] * F2[5] + F1[4] * F2[2];
  j.c = F1[2] * F2[5] - F1[4] * F2[3];
  j.d = F1[3] * F2[4] - F1[5] * F2[2];
  return j;
}
template <class R>
__device__ __forceinline__ cx<R> FFV1_0(const W<R>& F1, const W<R>& F2, const W<R>& V3, cx<R> COUP) {
  const LC<R> v = lightcone(V3);
  if constexpr (StableForms<R>::value) {
    const Bilinear<R> j = bilinear(F1, F2);
    const cx<R> TMP0 = j.a * v.a + j.b * v.b + j.c * v.c + j.d * v.d;
    return mul_mi(COUP * TMP0);
  } else {
    const cx<R> TMP0 = F1[2] * (F2[4] * v.a + F2[5] * v.c) + F1[3] * (F2[4] * v.d + F2[5] * v.b) +
                       F1[4] * (F2[2] * v.b - F2[3] * v.c) + F1[5] * (F2[3] * v.a - F2[2] * v.d);
    return mul_mi(COUP * TMP0);
  }
}

// FFV1_0 of two amplitudes that differ only in F2 (two helicity rows of one pass): the
// vertex is linear in F2, so the row vector L = F1 gamma.V is formed once and dotted with
// both spinors, TMP0 = sum_k L_k F2[k] (the same sum as FFV1_0, regrouped).
template <class R>
__device__ __forceinline__ void FFV1_0_pair2(const W<R>& F1, const W<R>& F2a, const W<R>& F2b, const W<R>& V3,
                                             cx<R> COUP, cx<R>& ampa, cx<R>& ampb) {
  const LC<R> v = lightcone(V3);
  const cx<R> L2 = F1[4] * v.b - F1[5] * v.d, L3 = F1[5] * v.a - F1[4] * v.c;
  const cx<R> L4 = F1[2] * v.a + F1[3] * v.d, L5 = F1[2] * v.c + F1[3] * v.b;
  ampa = mul_mi(COUP * (L2 * F2a[2] + L3 * F2a[3] + L4 * F2a[4] + L5 * F2a[5]));
  ampb = mul_mi(COUP * (L2 * F2b[2] + L3 * F2b[3] + L4 * F2b[4] + L5 * F2b[5]));
}
// The same for two amplitudes that differ only in F1: the column vector gamma.V F2 is shared.
template <class R>
__device__ __forceinline__ void FFV1_0_pair1(const W<R>& F1a, const W<R>& F1b, const W<R>& F2, const W<R>& V3,
                                             cx<R> COUP, cx<R>& ampa, cx<R>& ampb) {
  const LC<R> v = lightcone(V3);
  const cx<R> R2 = F2[4] * v.a + F2[5] * v.c, R3 = F2[4] * v.d + F2[5] * v.b;
  const cx<R> R4 = F2[2] * v.b - F2[3] * v.c, R5 = F2[3] * v.a - F2[2] * v.d;
  ampa = mul_mi(COUP * (F1a[2] * R2 + F1a[3] * R3 + F1a[4] * R4 + F1a[5] * R5));
  ampb = mul_mi(COUP * (F1b[2] * R2 + F1b[3] * R3 + F1b[4] * R4 + F1b[5] * R5));
}

// The same for two amplitudes that differ only in V: the fermion bilinear is shared.
template <class R>
__device__ __forceinline__ void FFV1_0_pairV(const W<R>& F1, const W<R>& F2, const W<R>& Va, const W<R>& Vb,
                                             cx<R> COUP, cx<R>& ampa, cx<R>& ampb) {
  const Bilinear<R> j = bilinear(F1, F2);
  const LC<R> a = lightcone(Va), b = lightcone(Vb);
  ampa = mul_mi(COUP * (j.a * a.a + j.b * a.b + j.c * a.c + j.d * a.d));
  ampb = mul_mi(COUP * (j.a * b.a + j.b * b.b + j.c * b.c + j.d * b.d));
}

// gamma.V F2 as the column (c2, c3, c4, c5) that FFV1_0 dots with F1 (the shared vector of
// FFV1_0_pair1) and that FFV1_1 multiplies by its propagator numerator.  A sum of
// FFV1_0 or FFV1_1 terms over different (F2, V) pairs with the same F1, or with the same
// off-shell momentum, is linear in this column: summing the columns first makes the
// dot product, or the propagator, a single step.
template <class R> struct VslashF { cx<R> c2, c3, c4, c5; };
template <class R>
__device__ __forceinline__ VslashF<R> vslash_f(const W<R>& F2, const W<R>& V3) {
  const LC<R> v = lightcone(V3);
  VslashF<R> x;
  x.c2 = F2[4] * v.a + F2[5] * v.c;
  x.c3 = F2[4] * v.d + F2[5] * v.b;
  x.c4 = F2[2] * v.b - F2[3] * v.c;
  x.c5 = F2[3] * v.a - F2[2] * v.d;
  return x;
}
template <class R> __device__ __forceinline__ void add_col(VslashF<R>& acc, const VslashF<R>& x) {
  acc.c2 += x.c2;
  acc.c3 += x.c3;
  acc.c4 += x.c4;
  acc.c5 += x.c5;
}
// FFV1_0(F1, F2, V3) summed over the (F2, V3) pairs whose columns make up x
template <class R> __device__ __forceinline__ cx<R> ffv1_0_col(const W<R>& F1, const VslashF<R>& x, cx<R> COUP) {
  return mul_mi(COUP * (F1[2] * x.c2 + F1[3] * x.c3 + F1[4] * x.c4 + F1[5] * x.c5));
}
// FFV1_1 spinor components (F1[2..5]) summed over the (F2, V3) pairs whose columns make
// up x, all with off-shell momentum P (slot convention) and den2 = P^2 - M1^2: the
// propagator (P.gamma + M1) / (den2 + i M1 W1) applied to the column
template <class R>
__device__ __forceinline__ void ffv1_1_col(const VslashF<R>& x, cx<R> COUP, R M1, R W1, R den2, const R (&P)[4],
                                           W<R>& F1) {
  const cx<R> di = mul_i(prop_denom(COUP, den2, M1, W1));
  const R pp0 = P[0] + P[3], pm0 = P[0] - P[3];
  const cx<R> q(P[1], P[2]), qb(P[1], -P[2]);
  F1[2] = di * (x.c2 * M1 - q * x.c5 - x.c4 * pp0);
  F1[3] = -di * (qb * x.c4 + x.c5 * pm0 - x.c3 * M1);
  F1[4] = -di * (x.c2 * pm0 - q * x.c3 - x.c4 * M1);
  F1[5] = di * (qb * x.c2 - x.c3 * pp0 + x.c5 * M1);
}

// ------------------------------------------------------------------------ FFV1_1
// P: the off-shell fermion's momentum (minus the sum of the input slots); the slots of F1
// are not written (callers that pass P from the external momenta never read them).
template <class R>
__device__ __forceinline__ void FFV1_1_p(const W<R>& F2, const W<R>& V3, cx<R> COUP, R M1, R W1, R den2,
                                         const R (&P)[4], W<R>& F1) {
  const cx<R> denom = prop_denom(COUP, den2, M1, W1);
  const LC<R> v = lightcone(V3);
  const R pp0 = P[0] + P[3], pm0 = P[0] - P[3];
  const cx<R> q(P[1], P[2]), qb(P[1], -P[2]);
  const cx<R> di = mul_i(denom);
  F1[2] = di * (F2[2] * (v.d * q - v.b * pp0) + F2[3] * (v.c * pp0 - v.a * q) + (F2[4] * v.a + F2[5] * v.c) * M1);
  F1[3] = -di * (F2[2] * (v.b * qb - v.d * pm0) + F2[3] * (v.a * pm0 - v.c * qb) - (F2[4] * v.d + F2[5] * v.b) * M1);
  F1[4] = -di * (F2[4] * (v.a * pm0 - v.d * q) + F2[5] * (v.c * pm0 - v.b * q) + (F2[3] * v.c - F2[2] * v.b) * M1);
  F1[5] = di * (F2[4] * (v.a * qb - v.d * pp0) + F2[5] * (v.c * qb - v.b * pp0) + (F2[3] * v.a - F2[2] * v.d) * M1);
}
template <class R>
__device__ __forceinline__ void FFV1_1(const W<R>& F2, const W<R>& V3, cx<R> COUP, R M1, R W1, R den2, W<R>& F1) {
  F1[0] = F2[0] + V3[0];
  F1[1] = F2[1] + V3[1];
  R P[4];
  mom_out(F1, P);
  FFV1_1_p(F2, V3, COUP, M1, W1, den2, P, F1);
}

// ------------------------------------------------------------------------ FFV1_2
template <class R>
__device__ __forceinline__ void FFV1_2_p(const W<R>& F1, const W<R>& V3, cx<R> COUP, R M2, R W2, R den2,
                                         const R (&P)[4], W<R>& F2) {
  const cx<R> denom = prop_denom(COUP, den2, M2, W2);
  const LC<R> v = lightcone(V3);
  const R pp0 = P[0] + P[3], pm0 = P[0] - P[3];
  const cx<R> q(P[1], P[2]), qb(P[1], -P[2]);
  const cx<R> di = mul_i(denom);
  F2[2] = di * (F1[2] * (v.a * pm0 - v.c * qb) + F1[3] * (v.d * pm0 - v.b * qb) + (F1[4] * v.b - F1[5] * v.d) * M2);
  F2[3] = -di * (F1[2] * (v.a * q - v.c * pp0) + F1[3] * (v.d * q - v.b * pp0) + (F1[4] * v.c - F1[5] * v.a) * M2);
  F2[4] = -di * (F1[4] * (v.c * qb - v.b * pp0) + F1[5] * (v.d * pp0 - v.a * qb) - (F1[2] * v.a + F1[3] * v.d) * M2);
  F2[5] = di * (F1[4] * (v.b * q - v.c * pm0) + F1[5] * (v.a * pm0 - v.d * q) + (F1[2] * v.c + F1[3] * v.b) * M2);
}
template <class R>
__device__ __forceinline__ void FFV1_2(const W<R>& F1, const W<R>& V3, cx<R> COUP, R M2, R W2, R den2, W<R>& F2) {
  F2[0] = F1[0] + V3[0];
  F2[1] = F1[1] + V3[1];
  R P[4];
  mom_out(F2, P);
  FFV1_2_p(F1, V3, COUP, M2, W2, den2, P, F2);
}

// ---------------------------------------------------------------------- VVV1P0_1
// P2, P3: the momenta of V2, V3 (their slot momenta); V1's slots are not written.
template <class R>
__device__ __forceinline__ void VVV1P0_1_p(const W<R>& V2, const W<R>& V3, cx<R> COUP, R M1, R W1, R den2,
                                           const R (&P2)[4], const R (&P3)[4], W<R>& V1) {
  R P1[4];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) P1[mu] = -(P2[mu] + P3[mu]);
  const cx<R> TMP1 = mdotp(V3, P1), TMP2 = mdotp(V3, P2), TMP3 = mdotp(V2, P1), TMP4 = mdotp(V2, P3);
  const cx<R> TMP5 = mdot(V3, V2);
  const cx<R> di = mul_i(prop_denom(COUP, den2, M1, W1));
  const cx<R> c2 = TMP2 - TMP1, c3 = TMP3 - TMP4;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) V1[2 + mu] = di * (TMP5 * (P3[mu] - P2[mu]) + V2[2 + mu] * c2 + V3[2 + mu] * c3);
}
template <class R>
__device__ __forceinline__ void VVV1P0_1(const W<R>& V2, const W<R>& V3, cx<R> COUP, R M1, R W1, R den2, W<R>& V1) {
  R P2[4], P3[4];
  mom_in(V2, P2);
  mom_in(V3, P3);
  V1[0] = V2[0] + V3[0];
  V1[1] = V2[1] + V3[1];
  VVV1P0_1_p(V2, V3, COUP, M1, W1, den2, P2, P3, V1);
}

// ------------------------------------------------------------------ VVVV{1,3,4}P0_1
// which = 1: L = V4 (V2.V3) - V3 (V2.V4); 3: V4 (V2.V3) - V2 (V3.V4); 4: V3 (V2.V4) - V2 (V3.V4)
template <int WHICH, class R>
__device__ __forceinline__ void VVVVP0_1(const W<R>& V2, const W<R>& V3, const W<R>& V4, cx<R> COUP, R M1, R W1,
                                         R den2, W<R>& V1) {
  V1[0] = V2[0] + V3[0] + V4[0];
  V1[1] = V2[1] + V3[1] + V4[1];
  const cx<R> dmi = mul_mi(prop_denom(COUP, den2, M1, W1));
  if (WHICH == 1) {
    const cx<R> x = dmi * mdot(V2, V3), y = dmi * mdot(V2, V4);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) V1[2 + mu] = V4[2 + mu] * x - V3[2 + mu] * y;
  } else if (WHICH == 3) {
    const cx<R> x = dmi * mdot(V2, V3), y = dmi * mdot(V3, V4);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) V1[2 + mu] = V4[2 + mu] * x - V2[2 + mu] * y;
  } else {
    const cx<R> x = dmi * mdot(V2, V4), y = dmi * mdot(V3, V4);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) V1[2 + mu] = V3[2 + mu] * x - V2[2 + mu] * y;
  }
}
template <class R>
__device__ __forceinline__ void VVVV1P0_1(const W<R>& a, const W<R>& b, const W<R>& c, cx<R> C, R M, R Wd, R den2,
                                          W<R>& o) {
  VVVVP0_1<1>(a, b, c, C, M, Wd, den2, o);
}
template <class R>
__device__ __forceinline__ void VVVV3P0_1(const W<R>& a, const W<R>& b, const W<R>& c, cx<R> C, R M, R Wd, R den2,
                                          W<R>& o) {
  VVVVP0_1<3>(a, b, c, C, M, Wd, den2, o);
}
template <class R>
__device__ __forceinline__ void VVVV4P0_1(const W<R>& a, const W<R>& b, const W<R>& c, cx<R> C, R M, R Wd, R den2,
                                          W<R>& o) {
  VVVVP0_1<4>(a, b, c, C, M, Wd, den2, o);
}

// ---------------------------------------------------------------------- FFV1P0_3
// den2 = P^2 - M3^2 of the off-shell vector, supplied by the caller
template <class R>
__device__ __forceinline__ void FFV1P0_3(const W<R>& F1, const W<R>& F2, cx<R> COUP, R M3, R W3, R den2, W<R>& V3) {
  V3[0] = F1[0] + F2[0];
  V3[1] = F1[1] + F2[1];
  const cx<R> dmi = mul_mi(prop_denom(COUP, den2, M3, W3));
  const cx<R> J0 = F1[2] * F2[4] + F1[3] * F2[5] + F1[4] * F2[2] + F1[5] * F2[3];
  const cx<R> J1 = F1[4] * F2[3] + F1[5] * F2[2] - F1[2] * F2[5] - F1[3] * F2[4];
  const cx<R> J2 = mul_i(F1[3] * F2[4] + F1[4] * F2[3] - F1[2] * F2[5] - F1[5] * F2[2]);
  const cx<R> J3 = F1[3] * F2[5] + F1[4] * F2[2] - F1[2] * F2[4] - F1[5] * F2[3];
  V3[2] = dmi * J0;
  V3[3] = dmi * J1;
  V3[4] = dmi * J2;
  V3[5] = dmi * J3;
}
template <class R>
__device__ __forceinline__ void FFV1P0_3(const W<R>& F1, const W<R>& F2, cx<R> COUP, R M3, R W3, W<R>& V3) {
  V3[0] = F1[0] + F2[0];
  V3[1] = F1[1] + F2[1];
  R P3[4];
  mom_out(V3, P3);
  const cx<R> dmi = mul_mi(prop_denom(COUP, den_from_slots(P3, M3), M3, W3));
  const cx<R> J0 = F1[2] * F2[4] + F1[3] * F2[5] + F1[4] * F2[2] + F1[5] * F2[3];
  const cx<R> J1 = F1[4] * F2[3] + F1[5] * F2[2] - F1[2] * F2[5] - F1[3] * F2[4];
  const cx<R> J2 = mul_i(F1[3] * F2[4] + F1[4] * F2[3] - F1[2] * F2[5] - F1[5] * F2[2]);
  const cx<R> J3 = F1[3] * F2[5] + F1[4] * F2[2] - F1[2] * F2[4] - F1[5] * F2[3];
  V3[2] = dmi * J0;
  V3[3] = dmi * J1;
  V3[4] = dmi * J2;
  V3[5] = dmi * J3;
}

// ------------------------------------------------------------------------ VVV1_0
template <class R>
__device__ __forceinline__ cx<R> VVV1_0(const W<R>& V1, const W<R>& V2, const W<R>& V3, cx<R> COUP) {
  R P1[4], P2[4], P3[4];
  mom_in(V1, P1);
  mom_in(V2, P2);
  mom_in(V3, P3);
  const cx<R> T = mdot(V1, V2) * (mdotp(V3, P1) - mdotp(V3, P2)) + mdot(V1, V3) * (mdotp(V2, P3) - mdotp(V2, P1)) +
                  mdot(V2, V3) * (mdotp(V1, P2) - mdotp(V1, P3));
  return mul_mi(COUP * T);
}

// ---------------------------------------------------------------- VVVV{1,3,4}_0
template <int WHICH, class R>
__device__ __forceinline__ cx<R> VVVV_0(const W<R>& V1, const W<R>& V2, const W<R>& V3, const W<R>& V4, cx<R> COUP) {
  cx<R> T;
  if (WHICH == 1) T = mdot(V1, V4) * mdot(V2, V3) - mdot(V1, V3) * mdot(V2, V4);
  else if (WHICH == 3) T = mdot(V1, V4) * mdot(V2, V3) - mdot(V1, V2) * mdot(V3, V4);
  else T = mdot(V1, V3) * mdot(V2, V4) - mdot(V1, V2) * mdot(V3, V4);
  return mul_mi(COUP * T);
}
template <class R>
__device__ __forceinline__ cx<R> VVVV1_0(const W<R>& a, const W<R>& b, const W<R>& c, const W<R>& d, cx<R> C) {
  return VVVV_0<1>(a, b, c, d, C);
}
template <class R>
__device__ __forceinline__ cx<R> VVVV3_0(const W<R>& a, const W<R>& b, const W<R>& c, const W<R>& d, cx<R> C) {
  return VVVV_0<3>(a, b, c, d, C);
}
template <class R>
__device__ __forceinline__ cx<R> VVVV4_0(const W<R>& a, const W<R>& b, const W<R>& c, const W<R>& d, cx<R> C) {
  return VVVV_0<4>(a, b, c, d, C);
}

// Slot-momentum variants of the off-shell routines (P^2 - M^2 from the current's own
// momentum, as the reference computes it); used by the building-block C entry points.
template <class R>
__device__ __forceinline__ void FFV1_1(const W<R>& F2, const W<R>& V3, cx<R> C, R M, R Wd, W<R>& F1) {
  R P[4] = {-(F2[0].re + V3[0].re), -(F2[1].re + V3[1].re), -(F2[1].im + V3[1].im), -(F2[0].im + V3[0].im)};
  FFV1_1(F2, V3, C, M, Wd, den_from_slots(P, M), F1);
}
template <class R>
__device__ __forceinline__ void FFV1_2(const W<R>& F1, const W<R>& V3, cx<R> C, R M, R Wd, W<R>& F2) {
  R P[4] = {-(F1[0].re + V3[0].re), -(F1[1].re + V3[1].re), -(F1[1].im + V3[1].im), -(F1[0].im + V3[0].im)};
  FFV1_2(F1, V3, C, M, Wd, den_from_slots(P, M), F2);
}
template <class R>
__device__ __forceinline__ void VVV1P0_1(const W<R>& V2, const W<R>& V3, cx<R> C, R M, R Wd, W<R>& V1) {
  R P[4] = {-(V2[0].re + V3[0].re), -(V2[1].re + V3[1].re), -(V2[1].im + V3[1].im), -(V2[0].im + V3[0].im)};
  VVV1P0_1(V2, V3, C, M, Wd, den_from_slots(P, M), V1);
}
template <int WHICH, class R>
__device__ __forceinline__ void VVVVP0_1s(const W<R>& a, const W<R>& b, const W<R>& c, cx<R> C, R M, R Wd, W<R>& o) {
  R P[4] = {-(a[0].re + b[0].re + c[0].re), -(a[1].re + b[1].re + c[1].re), -(a[1].im + b[1].im + c[1].im),
            -(a[0].im + b[0].im + c[0].im)};
  VVVVP0_1<WHICH>(a, b, c, C, M, Wd, den_from_slots(P, M), o);
}

}  // namespace mfx
