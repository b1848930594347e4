// ORACLE — test infrastructure only (see oracle_helas.hpp).
// Interpreter for generated HELAS programs: the CPU restatement of
//   madgraph_plugin/template_files/matrix_method_python.inc:80-138
//     smatrix: sum over all helicity rows of matrix(...), divided by `denominator`
//     matrix:  external wavefunctions -> HELAS call sequence -> jamp -> colour sum
// plus an independent hand restatement of the committed g g > t t~ module
// (python_package/madflow/tests/mockup_debug_me.py:419-541) that pins the
// generated g g > t t~ program.
#pragma once
#include "oracle_helas.hpp"
#include <vector>
#include <cstring>

namespace oracle {

enum Routine {
  R_FFV1_0 = 0, R_FFV1_1 = 1, R_FFV1_2 = 2, R_VVV1P0_1 = 3,
  R_VVVV1P0_1 = 4, R_VVVV3P0_1 = 5, R_VVVV4P0_1 = 6,
  R_FFV1P0_3 = 7, R_VVV1_0 = 8, R_VVVV1_0 = 9, R_VVVV3_0 = 10, R_VVVV4_0 = 11
};
enum External { E_V = 0, E_O = 1, E_I = 2 };

struct Proc {
  int nexternal, ncomb, nops, nwf, namps, ncolor, njterms;
  const double* helicities;  // ncomb * nexternal
  const int* externals;      // nexternal * 3: (kind, mass_idx, sign)
  const int* ops;            // nops * 9: (routine, out, a0, a1, a2, a3, coup, mass_idx, width_idx)
  const int* jterms;         // njterms * 2: (colour, amp)
  const double* jcoef;       // njterms * 2: (re, im) of the coefficient, real or imaginary
  const double* cf_num;      // ncolor * ncolor
  const double* cf_den;      // ncolor
  double denominator;
};

template <class T>
inline T param(const T* rp, int idx) { return idx < 0 ? T(0) : rp[idx]; }

// matrix() for one event and one helicity row: returns sum_ij J_i cf_ij conj(J_j)/den
// (real part), i.e. the colour-summed |M|^2 of this helicity before averaging.
template <class T>
T matrix_one(const Proc& P, const T* p /*nexternal*4*/, const T* hel, const T* rp, const cx<T>* coup,
             double sqh, std::vector<cx<T>>& wf, std::vector<cx<T>>& amp, std::vector<cx<T>>& jamp) {
  wf.assign((size_t)P.nwf * 6, cx<T>());
  amp.assign(P.namps, cx<T>());
  jamp.assign(P.ncolor, cx<T>());
  for (int l = 0; l < P.nexternal; ++l) {
    const int* e = P.externals + 3 * l;
    T mass = param(rp, e[1]);
    T sgn = T(e[2]);
    cx<T>* w = &wf[(size_t)l * 6];
    if (e[0] == E_V) vxxxxx(p + 4 * l, mass, hel[l], sgn, w, sqh);
    else if (e[0] == E_O) oxxxxx(p + 4 * l, mass, hel[l], sgn, w);
    else ixxxxx(p + 4 * l, mass, hel[l], sgn, w);
  }
  for (int k = 0; k < P.nops; ++k) {
    const int* o = P.ops + 9 * k;
    const cx<T>* a0 = &wf[(size_t)o[2] * 6];
    const cx<T>* a1 = o[3] >= 0 ? &wf[(size_t)o[3] * 6] : nullptr;
    const cx<T>* a2 = o[4] >= 0 ? &wf[(size_t)o[4] * 6] : nullptr;
    const cx<T>* a3 = o[5] >= 0 ? &wf[(size_t)o[5] * 6] : nullptr;
    cx<T> C = coup[o[6]];
    T M = param(rp, o[7]), W = param(rp, o[8]);
    cx<T>* out = (o[0] == R_FFV1_0 || o[0] >= R_VVV1_0) ? nullptr : &wf[(size_t)o[1] * 6];
    switch (o[0]) {
      case R_FFV1_0: amp[o[1]] = FFV1_0(a0, a1, a2, C); break;
      case R_FFV1_1: FFV1_1(a0, a1, C, M, W, out); break;
      case R_FFV1_2: FFV1_2(a0, a1, C, M, W, out); break;
      case R_VVV1P0_1: VVV1P0_1(a0, a1, C, M, W, out); break;
      case R_VVVV1P0_1: VVVVP0_1(1, a0, a1, a2, C, M, W, out); break;
      case R_VVVV3P0_1: VVVVP0_1(3, a0, a1, a2, C, M, W, out); break;
      case R_VVVV4P0_1: VVVVP0_1(4, a0, a1, a2, C, M, W, out); break;
      case R_FFV1P0_3: FFV1P0_3(a0, a1, C, M, W, out); break;
      case R_VVV1_0: amp[o[1]] = VVV1_0(a0, a1, a2, C); break;
      case R_VVVV1_0: amp[o[1]] = VVVV_0(1, a0, a1, a2, a3, C); break;
      case R_VVVV3_0: amp[o[1]] = VVVV_0(3, a0, a1, a2, a3, C); break;
      case R_VVVV4_0: amp[o[1]] = VVVV_0(4, a0, a1, a2, a3, C); break;
      default: break;
    }
  }
  // jamp lines (PyOut_exporter.py:334-375): coefficients +-1 or +-i (one quark line), or
  // rationals times 1 or i (two quark lines: the Fierz factors 1/2 and -1/6)
  for (int t = 0; t < P.njterms; ++t) {
    const int* j = P.jterms + 2 * t;
    const double re = P.jcoef[2 * t], im = P.jcoef[2 * t + 1];
    const cx<T>& a = amp[j[1]];
    cx<T> term;
    if (im == 0.0) term = (re == 1.0) ? a : (re == -1.0) ? -a : cx<T>(T(re) * a.re, T(re) * a.im);
    else if (im == 1.0) term = cx<T>(-a.im, a.re);
    else if (im == -1.0) term = cx<T>(a.im, -a.re);
    else term = cx<T>(-(T(im) * a.im), T(im) * a.re);
    jamp[j[0]] = jamp[j[0]] + term;
  }
  // colour sum (matrix_method_python.inc:137): sum_ij J_i cf_ij conj(J_j) / denom.
  // cf is real symmetric and every row of these processes has the same denominator,
  // so this equals sum_i Re(conj(J_i) * (sum_j cf_ij J_j)) / denom_i.
  T res = T(0);
  for (int i = 0; i < P.ncolor; ++i) {
    cx<T> z;
    for (int jj = 0; jj < P.ncolor; ++jj) {
      T c = T(P.cf_num[i * P.ncolor + jj]);
      z = cx<T>(z.re + c * jamp[jj].re, z.im + c * jamp[jj].im);
    }
    res = res + (z.re * jamp[i].re + z.im * jamp[i].im) / T(P.cf_den[i]);
  }
  return res;
}

template <class T>
T smatrix_one(const Proc& P, const T* p, const T* rp, const cx<T>* coup, double sqh,
              std::vector<cx<T>>& wf, std::vector<cx<T>>& amp, std::vector<cx<T>>& jamp) {
  T ans = T(0);
  std::vector<T> hel(P.nexternal);
  for (int h = 0; h < P.ncomb; ++h) {
    for (int l = 0; l < P.nexternal; ++l) hel[l] = T(P.helicities[h * P.nexternal + l]);
    ans = ans + matrix_one(P, p, hel.data(), rp, coup, sqh, wf, amp, jamp);
  }
  return ans / T(P.denominator);
}

// Hand restatement of Matrix_1_gg_ttx.matrix (mockup_debug_me.py:482-541).
// Params: mdl_MT, mdl_WT, GC_10, GC_11.
inline double matrix_gg_ttx_mockup(const double* all_ps, const double* hel, double MT, double WT,
                                   cx<double> GC_10, cx<double> GC_11, double sqh, cx<double>* jamp_out) {
  typedef cx<double> C;
  C w0[6], w1[6], w2[6], w3[6], w4[6];
  const double ZERO = 0.0;
  vxxxxx(all_ps + 0, ZERO, hel[0], -1.0, w0, sqh);
  vxxxxx(all_ps + 4, ZERO, hel[1], -1.0, w1, sqh);
  oxxxxx(all_ps + 8, MT, hel[2], +1.0, w2);
  ixxxxx(all_ps + 12, MT, hel[3], -1.0, w3);
  VVV1P0_1(w0, w1, GC_10, ZERO, ZERO, w4);
  C amp0 = FFV1_0(w3, w2, w4, GC_11);
  FFV1_1(w2, w0, GC_11, MT, WT, w4);
  C amp1 = FFV1_0(w3, w4, w1, GC_11);
  FFV1_2(w3, w0, GC_11, MT, WT, w4);
  C amp2 = FFV1_0(w4, w2, w1, GC_11);
  const C cI(0.0, 1.0);
  C jamp[2] = {cI * amp0 - amp1, -cI * amp0 - amp2};
  const double denom[2] = {3, 3};
  const double cf[2][2] = {{16, -2}, {-2, 16}};
  C matrix(0.0, 0.0);
  for (int i = 0; i < 2; ++i) {
    C ztemp(0.0, 0.0);
    for (int j = 0; j < 2; ++j) ztemp = ztemp + cf[i][j] * jamp[j];
    matrix = matrix + ztemp * conj(jamp[i]) / denom[i];
  }
  if (jamp_out) {
    jamp_out[0] = jamp[0];
    jamp_out[1] = jamp[1];
  }
  return matrix.re;
}

}  // namespace oracle
