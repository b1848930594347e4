// ORACLE — test infrastructure only.  CPU restatement of madflow's per-event
// wavefunctions and ALOHA routines, used by tests/, the smoke check and the CPU
// baseline of bench.py as the checker.  Never linked into the product library.
//
// Restates (reference paths relative to the madflow repository):
//   python_package/madflow/wavefunctions_flow.py   ixxxxx/oxxxxx/vxxxxx/sxxxxx
//   python_package/madflow/tests/mockup_debug_me.py FFV1_0/FFV1_1/FFV1_2/VVV1P0_1
//   madgraph_plugin/PyOut_create_aloha.py:124-328   emission rules (momentum from
//       slots [0],[1]; off-shell denominators; 'denom * (-i) * L' for vectors) used
//       for the routines the reference does not ship: FFV1P0_3, VVV1_0,
//       VVVV{1,3,4}_0, VVVV{1,3,4}P0_1 (Lorentz structures of the SM UFO).
//
// Everything is templated on the real type T so that the same code can run with
// an operation-counting scalar (flop counts per matrix element).
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

// ---------------------------------------------------------------- scalar helpers
template <class T> inline T osqrt(const T& x) { using std::sqrt; return sqrt(x); }
template <class T> inline T omin(const T& a, const T& b) { return (b < a) ? b : a; }
template <class T> inline T omax(const T& a, const T& b) { return (a < b) ? b : a; }
template <class T> inline T oabs(const T& a) { return (a < T(0)) ? -a : a; }
// tf.math.sign semantics: sign(0) = 0  (wavefunctions_flow.py:18-29)
template <class T> inline T tfsign(const T& y) { return (y > T(0)) ? T(1) : ((y < T(0)) ? T(-1) : T(0)); }

template <class T> struct cx {
  T re, im;
  cx() : re(0), im(0) {}
  cx(T r) : re(r), im(0) {}
  cx(T r, T i) : re(r), im(i) {}
};
template <class T> inline cx<T> operator+(const cx<T>& a, const cx<T>& b) { return cx<T>(a.re + b.re, a.im + b.im); }
template <class T> inline cx<T> operator-(const cx<T>& a, const cx<T>& b) { return cx<T>(a.re - b.re, a.im - b.im); }
template <class T> inline cx<T> operator-(const cx<T>& a) { return cx<T>(-a.re, -a.im); }
template <class T> inline cx<T> operator*(const cx<T>& a, const cx<T>& b) {
  return cx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T> inline cx<T> operator*(const cx<T>& a, const T& s) { return cx<T>(a.re * s, a.im * s); }
template <class T> inline cx<T> operator*(const T& s, const cx<T>& a) { return cx<T>(a.re * s, a.im * s); }
template <class T> inline cx<T> operator/(const cx<T>& a, const cx<T>& b) {
  T d = b.re * b.re + b.im * b.im;
  return cx<T>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
template <class T> inline cx<T> operator/(const cx<T>& a, const T& s) { return cx<T>(a.re / s, a.im / s); }
template <class T> inline cx<T> conj(const cx<T>& a) { return cx<T>(a.re, -a.im); }

// SQH: the reference computes float_me(sqrt(0.5)); Python floats pass through
// float32 first, so the value is 0.70710676908493042 (wavefunctions_flow.py:10).
struct Consts {
  double sqh;
};
inline Consts consts(bool ref_compat) {
  Consts c;
  c.sqh = ref_compat ? (double)(float)std::sqrt(0.5) : std::sqrt(0.5);
  return c;
}

// ------------------------------------------------------------ ixxxxx (:54-85)
template <class T>
void ixxxxx(const T p[4], T fmass, T nhel, T nsf, cx<T> fi[6]) {
  fi[0] = cx<T>(-p[0] * nsf, -p[3] * nsf);
  fi[1] = cx<T>(-p[1] * nsf, -p[2] * nsf);
  T nh = nhel * nsf;
  cx<T> v[4];
  if (fmass != T(0)) {
    // _ix_massive (:159-184)
    T pp = omin(p[0], osqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]));
    int ip = (int)std::floor((1.0 + (double)nh) / 2.0);
    int im = (int)std::floor((1.0 - (double)nh) / 2.0);
    if (pp == T(0)) {
      // _ox_massive_pp_zero(fmass, nsf, im, ip) with the projectors exchanged (:183, :365-388)
      int ipa = im, ima = ip;
      T sqm[2];
      sqm[0] = osqrt(oabs(fmass));
      sqm[1] = sqm[0] * tfsign(fmass);
      v[0] = cx<T>(T(ima) * sqm[std::abs(ima)], T(0));
      v[1] = cx<T>(T(ipa) * nsf * sqm[std::abs(ima)], T(0));
      v[2] = cx<T>(T(ima) * nsf * sqm[std::abs(ipa)], T(0));
      v[3] = cx<T>(T(ipa) * sqm[std::abs(ipa)], T(0));
    } else {
      // _ix_massive_pp_nonzero (:189-229)
      T sf[2] = {(T(1) + nsf + (T(1) - nsf) * nh) * T(0.5), (T(1) + nsf - (T(1) - nsf) * nh) * T(0.5)};
      T omega[2];
      omega[0] = osqrt(p[0] + pp);
      omega[1] = fmass / omega[0];
      T sfomeg[2] = {sf[0] * omega[ip], sf[1] * omega[im]};
      T pp3 = omax(pp + p[3], T(0));
      cx<T> chi1;
      if (pp3 == T(0)) {
        chi1 = cx<T>(-nh, T(0));
      } else {
        T d = osqrt(T(2) * pp * pp3);
        chi1 = cx<T>(nh * p[1] / d, p[2] / d);
      }
      cx<T> chi[2] = {cx<T>(osqrt(pp3 * T(0.5) / pp), T(0)), chi1};
      v[0] = cx<T>(sfomeg[0]) * chi[im];
      v[1] = cx<T>(sfomeg[0]) * chi[ip];
      v[2] = cx<T>(sfomeg[1]) * chi[im];
      v[3] = cx<T>(sfomeg[1]) * chi[ip];
    }
  } else {
    // _ix_massless (:234-256)
    T sqp0p3 = osqrt(omax(p[0] + p[3], T(0))) * nsf;
    cx<T> chi1;
    if (sqp0p3 == T(0)) chi1 = cx<T>(-nhel * osqrt(T(2) * p[0]), T(0));   // :261-275
    else chi1 = cx<T>(nh * p[1] / sqp0p3, p[2] / sqp0p3);               // :280-292
    cx<T> chi0(sqp0p3, T(0));
    if (nh == T(1)) {  // _ix_massless_nh_one (:296-311)
      v[0] = cx<T>(); v[1] = cx<T>(); v[2] = chi0; v[3] = chi1;
    } else {           // _ix_massless_nh_not_one (:315-330)
      v[0] = chi1; v[1] = chi0; v[2] = cx<T>(); v[3] = cx<T>();
    }
  }
  for (int k = 0; k < 4; ++k) fi[2 + k] = v[k];
}

// ------------------------------------------------------------ oxxxxx (:88-116)
template <class T>
void oxxxxx(const T p[4], T fmass, T nhel, T nsf, cx<T> fo[6]) {
  fo[0] = cx<T>(p[0] * nsf, p[3] * nsf);
  fo[1] = cx<T>(p[1] * nsf, p[2] * nsf);
  T nh = nhel * nsf;
  cx<T> v[4];
  if (fmass != T(0)) {
    // _ox_massive (:336-360)
    T pp = omin(p[0], osqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]));
    if (pp == T(0)) {
      // ip = -((1-nh)//2)*nhel, im = (1+nh)//2*nhel (:353-354)
      int ipa = (int)(-std::floor((1.0 - (double)nh) / 2.0) * (double)nhel);
      int ima = (int)(std::floor((1.0 + (double)nh) / 2.0) * (double)nhel);
      T sqm[2];
      sqm[0] = osqrt(oabs(fmass));
      sqm[1] = sqm[0] * tfsign(fmass);
      v[0] = cx<T>(T(ima) * sqm[std::abs(ima)], T(0));
      v[1] = cx<T>(T(ipa) * nsf * sqm[std::abs(ima)], T(0));
      v[2] = cx<T>(T(ima) * nsf * sqm[std::abs(ipa)], T(0));
      v[3] = cx<T>(T(ipa) * sqm[std::abs(ipa)], T(0));
    } else {
      // _ox_massive_pp_nonzero (:393-433)
      T sf[2] = {(T(1) + nsf + (T(1) - nsf) * nh) * T(0.5), (T(1) + nsf - (T(1) - nsf) * nh) * T(0.5)};
      T omega[2];
      omega[0] = osqrt(p[0] + pp);
      omega[1] = fmass / omega[0];
      int ip = (int)std::floor((1.0 + (double)nh) / 2.0);
      int im = (int)std::floor((1.0 - (double)nh) / 2.0);
      T sfomeg[2] = {sf[0] * omega[ip], sf[1] * omega[im]};
      T pp3 = omax(pp + p[3], T(0));
      cx<T> chi1;
      if (pp3 == T(0)) {
        chi1 = cx<T>(-nh, T(0));
      } else {
        T d = osqrt(T(2) * pp * pp3);
        chi1 = cx<T>(nh * p[1] / d, -p[2] / d);
      }
      cx<T> chi[2] = {cx<T>(osqrt(pp3 * T(0.5) / pp), T(0)), chi1};
      v[0] = cx<T>(sfomeg[1]) * chi[im];
      v[1] = cx<T>(sfomeg[1]) * chi[ip];
      v[2] = cx<T>(sfomeg[0]) * chi[im];
      v[3] = cx<T>(sfomeg[0]) * chi[ip];
    }
  } else {
    // _ox_massless (:438-462): p*mult with mult = (1,1,-1,1)
    T sqp0p3 = osqrt(omax(p[0] + p[3], T(0))) * nsf;
    cx<T> chi0;
    if (sqp0p3 == T(0)) chi0 = cx<T>(-nhel * osqrt(T(2) * p[0]), T(0));
    else chi0 = cx<T>(nh * p[1] / sqp0p3, -p[2] / sqp0p3);
    cx<T> chi1(sqp0p3, T(0));
    if (nh == T(1)) {  // _ix_massless_nh_not_one(chi) with chi = [chi0, sqp0p3]
      v[0] = chi1; v[1] = chi0; v[2] = cx<T>(); v[3] = cx<T>();
    } else {           // _ix_massless_nh_one(chi)
      v[0] = cx<T>(); v[1] = cx<T>(); v[2] = chi0; v[3] = chi1;
    }
  }
  for (int k = 0; k < 4; ++k) fo[2 + k] = v[k];
}

// ------------------------------------------------------------ vxxxxx (:119-154)
template <class T>
void vxxxxx(const T p[4], T vmass, T nhel, T nsv, cx<T> eps[6], double sqh_d) {
  const T SQH = T(sqh_d);
  eps[0] = cx<T>(p[0] * nsv, p[3] * nsv);
  eps[1] = cx<T>(p[1] * nsv, p[2] * nsv);
  T pt2 = p[1] * p[1] + p[2] * p[2];
  T pp = omin(p[0], osqrt(pt2 + p[3] * p[3]));
  T pt = omin(pp, osqrt(pt2));
  T hel0 = T(1) - oabs(nhel);
  T nsvahl = nsv * oabs(nhel);
  cx<T> v[4];
  if (nhel == T(4)) {
    // BRST check: epsilon = p/p0 (massless) or p/vmass (:469-515)
    T d = (vmass == T(0)) ? p[0] : vmass;
    for (int k = 0; k < 4; ++k) v[k] = cx<T>(p[k] / d, T(0));
  } else if (vmass != T(0)) {
    // _vx_no_BRST_check_massive (:548-571)
    if (pp == T(0)) {
      // :576-593 (v[0] keeps the complex 1 of the reference's initial list)
      v[0] = cx<T>(T(1), T(0));
      v[1] = cx<T>(-nhel * SQH, T(0));
      v[2] = cx<T>(T(0), nsvahl * SQH);
      v[3] = cx<T>(hel0, T(0));
    } else {
      // :597-626
      T emp = p[0] / (vmass * pp);
      v[0] = cx<T>(hel0 * pp / vmass, T(0));
      v[3] = cx<T>(hel0 * p[3] * emp + nhel * pt / pp * SQH, T(0));
      if (pt != T(0)) {
        T pzpt = p[3] / (pp * pt) * SQH * nhel;
        v[1] = cx<T>(hel0 * p[1] * emp - p[1] * pzpt, -nsvahl * p[2] / pt * SQH);
        v[2] = cx<T>(hel0 * p[2] * emp - p[2] * pzpt, nsvahl * p[1] / pt * SQH);
      } else {
        v[1] = cx<T>(-nhel * SQH, T(0));
        v[2] = cx<T>(T(0), nsvahl * SQH * tfsign(p[3]));
      }
    }
  } else {
    // _vx_no_BRST_check_massless (:683-747): pp is p0 here
    T ppl = p[0];
    T ptl = osqrt(p[1] * p[1] + p[2] * p[2]);
    v[0] = cx<T>();
    v[3] = cx<T>(nhel * ptl / ppl * SQH, T(0));
    if (ptl != T(0)) {
      T pzpt = p[3] / (ppl * ptl) * SQH * nhel;
      v[1] = cx<T>(-p[1] * pzpt, -nsv * p[2] / ptl * SQH);
      v[2] = cx<T>(-p[2] * pzpt, nsv * p[1] / ptl * SQH);
    } else {
      v[1] = cx<T>(-nhel * SQH, T(0));
      v[2] = cx<T>(T(0), nsv * SQH * tfsign(p[3]));
    }
  }
  for (int k = 0; k < 4; ++k) eps[2 + k] = v[k];
}

// ------------------------------------------------------------ sxxxxx (:32-51)
template <class T>
void sxxxxx(const T p[4], T nss, cx<T> phi[3]) {
  phi[0] = cx<T>(p[0] * nss, p[3] * nss);
  phi[1] = cx<T>(p[1] * nss, p[2] * nss);
  phi[2] = cx<T>(T(1), T(0));
}

// =============================================================== ALOHA routines
// Multiplication by +-i is a component swap (no arithmetic); the reference writes it
// as a complex product with cI = complex_tf(0,1), which gives identical IEEE results.
struct ImagUnit { bool neg; };
constexpr ImagUnit cI_{false};
inline ImagUnit operator-(ImagUnit u) { return ImagUnit{!u.neg}; }
template <class T> inline cx<T> operator*(ImagUnit u, const cx<T>& a) {
  return u.neg ? cx<T>(a.im, -a.re) : cx<T>(-a.im, a.re);
}
template <class T> inline cx<T> operator*(const cx<T>& a, ImagUnit u) { return u * a; }
template <class T> inline cx<T> CI() { return cx<T>(T(0), T(1)); }

// FFV1_0 (mockup_debug_me.py:34-54)
template <class T>
cx<T> FFV1_0(const cx<T>* F1, const cx<T>* F2, const cx<T>* V3, cx<T> COUP) {
  const ImagUnit cI = cI_;
  cx<T> TMP0 = F1[2] * (F2[4] * (V3[2] + V3[5]) + F2[5] * (V3[3] + cI * V3[4])) +
               (F1[3] * (F2[4] * (V3[3] - cI * V3[4]) + F2[5] * (V3[2] - V3[5])) +
                (F1[4] * (F2[2] * (V3[2] - V3[5]) - F2[3] * (V3[3] + cI * V3[4])) +
                 F1[5] * (F2[2] * (-V3[3] + cI * V3[4]) + F2[3] * (V3[2] + V3[5]))));
  return COUP * (-cI) * TMP0;
}

template <class T>
inline void momentum_in(const cx<T>* W, T P[4]) {  // P = (Re s0, Re s1, Im s1, Im s0)
  P[0] = W[0].re; P[1] = W[1].re; P[2] = W[1].im; P[3] = W[0].im;
}
template <class T>
inline void momentum_out(const cx<T>* W, T P[4]) {  // P = -(Re s0, Re s1, Im s1, Im s0)
  P[0] = -W[0].re; P[1] = -W[1].re; P[2] = -W[1].im; P[3] = -W[0].im;
}
template <class T>
inline cx<T> propagator_denom(cx<T> COUP, const T P[4], T M, T W) {
  // COUP / (P^2 - M*(M - i W)) = COUP / ((P^2 - M^2) + i M W)
  cx<T> d(P[0] * P[0] - P[1] * P[1] - P[2] * P[2] - P[3] * P[3] - M * M, M * W);
  return COUP / d;
}

// FFV1_1 (mockup_debug_me.py:57-194)
template <class T>
void FFV1_1(const cx<T>* F2, const cx<T>* V3, cx<T> COUP, T M1r, T W1r, cx<T>* F1) {
  const ImagUnit cI = cI_;
  const cx<T> M1(M1r);
  F1[0] = F2[0] + V3[0];
  F1[1] = F2[1] + V3[1];
  T Pr[4];
  momentum_out(F1, Pr);
  const T* P1 = Pr;
  cx<T> denom = propagator_denom(COUP, Pr, M1r, W1r);
    F1[2] = denom * cI *
          (F2[2] * (P1[0] * (-V3[2] + V3[5]) + (P1[1] * (V3[3] - cI * V3[4]) +
                                                 (P1[2] * (cI * V3[3] + V3[4]) + P1[3] * (-V3[2] + V3[5])))) +
           (F2[3] * (P1[0] * (V3[3] + cI * V3[4]) +
                     ((-P1[1]) * (V3[2] + V3[5]) + ((-P1[2]) * (cI * (V3[2] + V3[5])) + P1[3] * (V3[3] + cI * V3[4])))) +
            M1 * (F2[4] * (V3[2] + V3[5]) + F2[5] * (V3[3] + cI * V3[4]))));
  F1[3] = denom * (-cI) *
          (F2[2] * (P1[0] * (-V3[3] + cI * V3[4]) +
                    (P1[1] * (V3[2] - V3[5]) + (P1[2] * (-cI * V3[2] + cI * V3[5]) + P1[3] * (V3[3] - cI * V3[4])))) +
           (F2[3] * (P1[0] * (V3[2] + V3[5]) +
                     ((-P1[1]) * (V3[3] + cI * V3[4]) + (P1[2] * (cI * V3[3] - V3[4]) - P1[3] * (V3[2] + V3[5])))) +
            M1 * (F2[4] * (-V3[3] + cI * V3[4]) + F2[5] * (-V3[2] + V3[5]))));
  F1[4] = denom * (-cI) *
          (F2[4] * (P1[0] * (V3[2] + V3[5]) +
                    (P1[1] * (-V3[3] + cI * V3[4]) + ((-P1[2]) * (cI * V3[3] + V3[4]) - P1[3] * (V3[2] + V3[5])))) +
           (F2[5] * (P1[0] * (V3[3] + cI * V3[4]) +
                     (P1[1] * (-V3[2] + V3[5]) + (P1[2] * (-cI * V3[2] + cI * V3[5]) - P1[3] * (V3[3] + cI * V3[4])))) +
            M1 * (F2[2] * (-V3[2] + V3[5]) + F2[3] * (V3[3] + cI * V3[4]))));
  F1[5] = denom * cI *
          (F2[4] * (P1[0] * (-V3[3] + cI * V3[4]) +
                    (P1[1] * (V3[2] + V3[5]) + ((-P1[2]) * (cI * (V3[2] + V3[5])) + P1[3] * (-V3[3] + cI * V3[4])))) +
           (F2[5] * (P1[0] * (-V3[2] + V3[5]) +
                     (P1[1] * (V3[3] + cI * V3[4]) + (P1[2] * (-cI * V3[3] + V3[4]) + P1[3] * (-V3[2] + V3[5])))) +
            M1 * (F2[2] * (-V3[3] + cI * V3[4]) + F2[3] * (V3[2] + V3[5]))));
}

// FFV1_2 (mockup_debug_me.py:197-328)
template <class T>
void FFV1_2(const cx<T>* F1, const cx<T>* V3, cx<T> COUP, T M2r, T W2r, cx<T>* F2) {
  const ImagUnit cI = cI_;
  const cx<T> M2(M2r);
  F2[0] = F1[0] + V3[0];
  F2[1] = F1[1] + V3[1];
  T Pr[4];
  momentum_out(F2, Pr);
  const T* P2 = Pr;
  cx<T> denom = propagator_denom(COUP, Pr, M2r, W2r);
    F2[2] = denom * cI *
          (F1[2] * (P2[0] * (V3[2] + V3[5]) +
                    ((-P2[1]) * (V3[3] + cI * V3[4]) + (P2[2] * (cI * V3[3] - V3[4]) - P2[3] * (V3[2] + V3[5])))) +
           (F1[3] * (P2[0] * (V3[3] - cI * V3[4]) +
                     (P2[1] * (-V3[2] + V3[5]) + (P2[2] * (cI * V3[2] - cI * V3[5]) + P2[3] * (-V3[3] + cI * V3[4])))) +
            M2 * (F1[4] * (V3[2] - V3[5]) + F1[5] * (-V3[3] + cI * V3[4]))));
  F2[3] = denom * (-cI) *
          (F1[2] * ((-P2[0]) * (V3[3] + cI * V3[4]) +
                    (P2[1] * (V3[2] + V3[5]) + (P2[2] * (cI * (V3[2] + V3[5])) - P2[3] * (V3[3] + cI * V3[4])))) +
           (F1[3] * (P2[0] * (-V3[2] + V3[5]) +
                     (P2[1] * (V3[3] - cI * V3[4]) + (P2[2] * (cI * V3[3] + V3[4]) + P2[3] * (-V3[2] + V3[5])))) +
            M2 * (F1[4] * (V3[3] + cI * V3[4]) - F1[5] * (V3[2] + V3[5]))));
  F2[4] = denom * (-cI) *
          (F1[4] * (P2[0] * (-V3[2] + V3[5]) +
                    (P2[1] * (V3[3] + cI * V3[4]) + (P2[2] * (-cI * V3[3] + V3[4]) + P2[3] * (-V3[2] + V3[5])))) +
           (F1[5] * (P2[0] * (V3[3] - cI * V3[4]) +
                     ((-P2[1]) * (V3[2] + V3[5]) + (P2[2] * (cI * (V3[2] + V3[5])) + P2[3] * (V3[3] - cI * V3[4])))) +
            M2 * ((-F1[2]) * (V3[2] + V3[5]) + F1[3] * (-V3[3] + cI * V3[4]))));
  F2[5] = denom * cI *
          (F1[4] * ((-P2[0]) * (V3[3] + cI * V3[4]) +
                    (P2[1] * (V3[2] - V3[5]) + (P2[2] * (cI * V3[2] - cI * V3[5]) + P2[3] * (V3[3] + cI * V3[4])))) +
           (F1[5] * (P2[0] * (V3[2] + V3[5]) +
                     (P2[1] * (-V3[3] + cI * V3[4]) + ((-P2[2]) * (cI * V3[3] + V3[4]) - P2[3] * (V3[2] + V3[5])))) +
            M2 * (F1[2] * (V3[3] + cI * V3[4]) + F1[3] * (V3[2] - V3[5]))));
}

template <class T>
inline cx<T> mdot(const cx<T>* A, const cx<T>* B) {  // Minkowski dot of slots 2..5
  return A[2] * B[2] - A[3] * B[3] - A[4] * B[4] - A[5] * B[5];
}
template <class T>
inline cx<T> mdotp(const cx<T>* A, const T* P) {
  return A[2] * P[0] - A[3] * P[1] - A[4] * P[2] - A[5] * P[3];
}

// VVV1P0_1 (mockup_debug_me.py:331-397)
template <class T>
void VVV1P0_1(const cx<T>* V2, const cx<T>* V3, cx<T> COUP, T M1, T W1, cx<T>* V1) {
  const ImagUnit cI = cI_;
  T P2[4], P3[4], P1[4];
  momentum_in(V2, P2);
  momentum_in(V3, P3);
  V1[0] = V2[0] + V3[0];
  V1[1] = V2[1] + V3[1];
  momentum_out(V1, P1);
  cx<T> TMP1 = mdotp(V3, P1);
  cx<T> TMP2 = mdotp(V3, P2);
  cx<T> TMP3 = mdotp(V2, P1);
  cx<T> TMP4 = mdotp(V2, P3);
  cx<T> TMP5 = mdot(V3, V2);
  cx<T> denom = propagator_denom(COUP, P1, M1, W1);
  for (int mu = 0; mu < 4; ++mu) {
    V1[2 + mu] = denom * (cI * (TMP5 * (P3[mu] - P2[mu])) +
                          (V2[2 + mu] * (-cI * TMP1 + cI * TMP2) + V3[2 + mu] * (cI * TMP3 - cI * TMP4)));
  }
}

// ---------------------------------------------------- routines derived from the SM UFO
// FFV1P0_3: off-shell vector from a fermion pair, Lorentz Gamma(3,2,1).
// J^mu = F1bar gamma^mu F2 in the HELAS chiral basis; V3 = denom * (-i) * J.
template <class T>
void FFV1P0_3(const cx<T>* F1, const cx<T>* F2, cx<T> COUP, T M3, T W3, cx<T>* V3) {
  const ImagUnit cI = cI_;
  V3[0] = F1[0] + F2[0];
  V3[1] = F1[1] + F2[1];
  T P3[4];
  momentum_out(V3, P3);
  cx<T> denom = propagator_denom(COUP, P3, M3, W3);
  cx<T> J0 = F1[2] * F2[4] + F1[3] * F2[5] + F1[4] * F2[2] + F1[5] * F2[3];
  cx<T> J1 = -F1[2] * F2[5] - F1[3] * F2[4] + F1[4] * F2[3] + F1[5] * F2[2];
  cx<T> J2 = -cI * (F1[2] * F2[5] + F1[5] * F2[2]) + cI * (F1[3] * F2[4] + F1[4] * F2[3]);
  cx<T> J3 = -F1[2] * F2[4] - F1[5] * F2[3] + F1[3] * F2[5] + F1[4] * F2[2];
  V3[2] = denom * (-cI) * J0;
  V3[3] = denom * (-cI) * J1;
  V3[4] = denom * (-cI) * J2;
  V3[5] = denom * (-cI) * J3;
}

// VVV1_0: COUP * (-i) * VVV1(1,2,3) with VVV1 = g12 (P1-P2).e3 + g13 (P3-P1).e2 + g23 (P2-P3).e1
template <class T>
cx<T> VVV1_0(const cx<T>* V1, const cx<T>* V2, const cx<T>* V3, cx<T> COUP) {
  const ImagUnit cI = cI_;
  T P1[4], P2[4], P3[4];
  momentum_in(V1, P1);
  momentum_in(V2, P2);
  momentum_in(V3, P3);
  cx<T> TMP = mdot(V1, V2) * (mdotp(V3, P1) - mdotp(V3, P2)) +
              mdot(V1, V3) * (mdotp(V2, P3) - mdotp(V2, P1)) +
              mdot(V2, V3) * (mdotp(V1, P2) - mdotp(V1, P3));
  return COUP * (-cI) * TMP;
}

// VVVV1 = g14 g23 - g13 g24 ; VVVV3 = g14 g23 - g12 g34 ; VVVV4 = g13 g24 - g12 g34
template <class T>
cx<T> VVVV_0(int which, const cx<T>* V1, const cx<T>* V2, const cx<T>* V3, const cx<T>* V4, cx<T> COUP) {
  const ImagUnit cI = cI_;
  cx<T> TMP;
  if (which == 1) TMP = mdot(V1, V4) * mdot(V2, V3) - mdot(V1, V3) * mdot(V2, V4);
  else if (which == 3) TMP = mdot(V1, V4) * mdot(V2, V3) - mdot(V1, V2) * mdot(V3, V4);
  else TMP = mdot(V1, V3) * mdot(V2, V4) - mdot(V1, V2) * mdot(V3, V4);
  return COUP * (-cI) * TMP;
}

// VVVV{1,3,4}P0_1: leg 1 off shell, V1^mu = denom * (-i) * L^mu(V2,V3,V4)
template <class T>
void VVVVP0_1(int which, const cx<T>* V2, const cx<T>* V3, const cx<T>* V4, cx<T> COUP, T M1, T W1, cx<T>* V1) {
  const ImagUnit cI = cI_;
  V1[0] = V2[0] + V3[0] + V4[0];
  V1[1] = V2[1] + V3[1] + V4[1];
  T P1[4];
  momentum_out(V1, P1);
  cx<T> denom = propagator_denom(COUP, P1, M1, W1);
  cx<T> d23 = mdot(V2, V3), d24 = mdot(V2, V4), d34 = mdot(V3, V4);
  for (int mu = 0; mu < 4; ++mu) {
    cx<T> L;
    if (which == 1) L = V4[2 + mu] * d23 - V3[2 + mu] * d24;
    else if (which == 3) L = V4[2 + mu] * d23 - V2[2 + mu] * d34;
    else L = V3[2 + mu] * d24 - V2[2 + mu] * d34;
    V1[2 + mu] = denom * (-cI) * L;
  }
}

}  // namespace oracle
