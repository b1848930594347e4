// ORACLE — test infrastructure only.  extern "C" entry points of the CPU oracle,
// loaded with ctypes by tests/, the smoke check and bench.py's cpu_baseline leg.
// Nothing in the product library links or calls this file.
#include "oracle_process.hpp"
#include "oracle_phasespace.hpp"
#include <cstdio>
#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

// ----------------------------------------------------------- flop-counting scalar
struct Flop {
  double v;
  static thread_local uint64_t add, mul, div, sqrt_;
  Flop() : v(0) {}
  Flop(double x) : v(x) {}
  explicit operator double() const { return v; }
};
thread_local uint64_t Flop::add = 0, Flop::mul = 0, Flop::div = 0, Flop::sqrt_ = 0;
inline Flop operator+(Flop a, Flop b) { ++Flop::add; return Flop(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { ++Flop::add; return Flop(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { ++Flop::mul; return Flop(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { ++Flop::div; return Flop(a.v / b.v); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }
inline Flop sqrt(Flop a) { ++Flop::sqrt_; return Flop(std::sqrt(a.v)); }

static Proc make_proc(int nexternal, int ncomb, const double* hel, const int* ext, int nops, const int* ops, int nwf,
                      int namps, int ncolor, int njterms, const int* jterms, const double* jcoef,
                      const double* cf_num, const double* cf_den, double denominator) {
  Proc P;
  P.nexternal = nexternal; P.ncomb = ncomb; P.helicities = hel; P.externals = ext;
  P.nops = nops; P.ops = ops; P.nwf = nwf; P.namps = namps; P.ncolor = ncolor;
  P.njterms = njterms; P.jterms = jterms; P.jcoef = jcoef;
  P.cf_num = cf_num; P.cf_den = cf_den; P.denominator = denominator;
  return P;
}

#define PROC_ARGS                                                                                          \
  int nexternal, int ncomb, const double *hel, const int *ext, int nops, const int *ops, int nwf, int namps, \
      int ncolor, int njterms, const int *jterms, const double *jcoef, const double *cf_num,                 \
      const double *cf_den, double denominator
#define PROC_PASS \
  nexternal, ncomb, hel, ext, nops, ops, nwf, namps, ncolor, njterms, jterms, jcoef, cf_num, cf_den, denominator

extern "C" {

int mfo_version() { return 1; }

// smatrix for nev events.  rp: real params; coup: n_coup complex (re,im) values, each
// either one value (coup_len=1) or one per event (coup_len=nev).
int mfo_smatrix(PROC_ARGS, const double* all_ps, int64_t nev, const double* rp, const double* coup, int n_coup,
                int64_t coup_len, double* out, int ref_compat, int nthreads) {
  Proc P = make_proc(PROC_PASS);
  double sqh = consts(ref_compat != 0).sqh;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    std::vector<cx<double>> wf, amp, jamp, cp(n_coup);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t e = 0; e < nev; ++e) {
      for (int c = 0; c < n_coup; ++c) {
        int64_t k = coup_len == 1 ? 0 : e;
        cp[c] = cx<double>(coup[2 * (c * coup_len + k)], coup[2 * (c * coup_len + k) + 1]);
      }
      out[e] = smatrix_one(P, all_ps + e * nexternal * 4, rp, cp.data(), sqh, wf, amp, jamp);
    }
  }
  return 0;
}

// matrix() for one helicity vector (length nexternal, may contain 4 for a BRST check)
int mfo_matrix_hel(PROC_ARGS, const double* all_ps, int64_t nev, const double* helv, const double* rp,
                   const double* coup, int n_coup, double* out, double* jamp_out, int ref_compat) {
  Proc P = make_proc(PROC_PASS);
  double sqh = consts(ref_compat != 0).sqh;
  std::vector<cx<double>> wf, amp, jamp, cp(n_coup);
  for (int c = 0; c < n_coup; ++c) cp[c] = cx<double>(coup[2 * c], coup[2 * c + 1]);
  for (int64_t e = 0; e < nev; ++e) {
    out[e] = matrix_one(P, all_ps + e * nexternal * 4, helv, rp, cp.data(), sqh, wf, amp, jamp);
    if (jamp_out)
      for (int i = 0; i < ncolor; ++i) {
        jamp_out[(e * ncolor + i) * 2] = jamp[i].re;
        jamp_out[(e * ncolor + i) * 2 + 1] = jamp[i].im;
      }
  }
  return 0;
}

// Exact operation count of one smatrix evaluation (add/sub, mul, div, sqrt).
int mfo_flops(PROC_ARGS, const double* all_ps, const double* rp, const double* coup, int n_coup, uint64_t* counts) {
  Proc P = make_proc(PROC_PASS);
  std::vector<Flop> p(nexternal * 4), r(4);
  for (int i = 0; i < nexternal * 4; ++i) p[i] = Flop(all_ps[i]);
  for (int i = 0; i < 4; ++i) r[i] = Flop(rp[i]);
  std::vector<cx<Flop>> wf, amp, jamp, cp(n_coup);
  for (int c = 0; c < n_coup; ++c) cp[c] = cx<Flop>(Flop(coup[2 * c]), Flop(coup[2 * c + 1]));
  Flop::add = Flop::mul = Flop::div = Flop::sqrt_ = 0;
  Flop res = smatrix_one(P, p.data(), r.data(), cp.data(), consts(false).sqh, wf, amp, jamp);
  counts[0] = Flop::add; counts[1] = Flop::mul; counts[2] = Flop::div; counts[3] = Flop::sqrt_;
  (void)res;
  return 0;
}

int mfo_smatrix_gg_ttx_mockup(const double* all_ps, int64_t nev, double MT, double WT, const double* gc10,
                              const double* gc11, double* out, double* jamp_out, int ref_compat) {
  double sqh = consts(ref_compat != 0).sqh;
  static const double hels[16][4] = {
      {-1, -1, -1, 1}, {-1, -1, -1, -1}, {-1, -1, 1, 1}, {-1, -1, 1, -1}, {-1, 1, -1, 1}, {-1, 1, -1, -1},
      {-1, 1, 1, 1},   {-1, 1, 1, -1},   {1, -1, -1, 1}, {1, -1, -1, -1}, {1, -1, 1, 1},  {1, -1, 1, -1},
      {1, 1, -1, 1},   {1, 1, -1, -1},   {1, 1, 1, 1},   {1, 1, 1, -1}};
  cx<double> GC10(gc10[0], gc10[1]), GC11(gc11[0], gc11[1]);
  for (int64_t e = 0; e < nev; ++e) {
    double ans = 0;
    for (int h = 0; h < 16; ++h) {
      cx<double> jm[2];
      ans += matrix_gg_ttx_mockup(all_ps + 16 * e, hels[h], MT, WT, GC10, GC11, sqh, jm);
      if (jamp_out && h == 0) {
        jamp_out[4 * e + 0] = jm[0].re; jamp_out[4 * e + 1] = jm[0].im;
        jamp_out[4 * e + 2] = jm[1].re; jamp_out[4 * e + 3] = jm[1].im;
      }
    }
    out[e] = ans / 256.0;
  }
  return 0;
}

// Single wavefunction: kind 0 v, 1 o, 2 i, 3 s.  out: 6 complex (3 for s) per event, (re,im).
int mfo_wavefunction(int kind, const double* p, int64_t nev, double mass, double nhel, double nsign, double* out,
                     int ref_compat) {
  double sqh = consts(ref_compat != 0).sqh;
  for (int64_t e = 0; e < nev; ++e) {
    cx<double> w[6];
    if (kind == 0) vxxxxx(p + 4 * e, mass, nhel, nsign, w, sqh);
    else if (kind == 1) oxxxxx(p + 4 * e, mass, nhel, nsign, w);
    else if (kind == 2) ixxxxx(p + 4 * e, mass, nhel, nsign, w);
    else sxxxxx(p + 4 * e, nsign, w);
    int n = kind == 3 ? 3 : 6;
    for (int k = 0; k < n; ++k) {
      out[(e * n + k) * 2] = w[k].re;
      out[(e * n + k) * 2 + 1] = w[k].im;
    }
  }
  return 0;
}

// One ALOHA routine on nev sets of inputs.  ins: n_in wavefunctions of 6 complex per
// event laid out [input][event][6][2].  For amplitude routines out is 1 complex per
// event; for off-shell routines 6 complex per event.
int mfo_aloha(int routine, const double* ins, int n_in, int64_t nev, const double* coup, double M, double W,
              double* out) {
  for (int64_t e = 0; e < nev; ++e) {
    cx<double> in[4][6], o[6];
    for (int i = 0; i < n_in; ++i)
      for (int k = 0; k < 6; ++k) {
        const double* s = ins + (((size_t)i * nev + e) * 6 + k) * 2;
        in[i][k] = cx<double>(s[0], s[1]);
      }
    cx<double> C(coup[0], coup[1]);
    cx<double> amp;
    bool is_amp = false;
    switch (routine) {
      case R_FFV1_0: amp = FFV1_0(in[0], in[1], in[2], C); is_amp = true; break;
      case R_FFV1_1: FFV1_1(in[0], in[1], C, M, W, o); break;
      case R_FFV1_2: FFV1_2(in[0], in[1], C, M, W, o); break;
      case R_VVV1P0_1: VVV1P0_1(in[0], in[1], C, M, W, o); break;
      case R_VVVV1P0_1: VVVVP0_1(1, in[0], in[1], in[2], C, M, W, o); break;
      case R_VVVV3P0_1: VVVVP0_1(3, in[0], in[1], in[2], C, M, W, o); break;
      case R_VVVV4P0_1: VVVVP0_1(4, in[0], in[1], in[2], C, M, W, o); break;
      case R_FFV1P0_3: FFV1P0_3(in[0], in[1], C, M, W, o); break;
      case R_VVV1_0: amp = VVV1_0(in[0], in[1], in[2], C); is_amp = true; break;
      case R_VVVV1_0: amp = VVVV_0(1, in[0], in[1], in[2], in[3], C); is_amp = true; break;
      case R_VVVV3_0: amp = VVVV_0(3, in[0], in[1], in[2], in[3], C); is_amp = true; break;
      case R_VVVV4_0: amp = VVVV_0(4, in[0], in[1], in[2], in[3], C); is_amp = true; break;
      default: return -1;
    }
    if (is_amp) {
      out[2 * e] = amp.re;
      out[2 * e + 1] = amp.im;
    } else {
      for (int k = 0; k < 6; ++k) {
        out[(e * 6 + k) * 2] = o[k].re;
        out[(e * 6 + k) * 2 + 1] = o[k].im;
      }
    }
  }
  return 0;
}

int mfo_rambo(const double* xrand, int64_t nev, int xstride, int n, const double* sqrts, int sqrts_stride,
              const double* masses, int ref_compat, double* p_out, double* wt_out) {
  rambo_batch(xrand, nev, xstride, n, sqrts, sqrts_stride, masses, ref_compat != 0, p_out, wt_out);
  return 0;
}

int mfo_philox_uniforms(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, double* out) {
  for (int64_t e = 0; e < nev; ++e) philox_uniforms(ev0 + e, ndim, seed, iteration, out + e * ndim);
  return 0;
}

// Hit-or-miss unweighting (the device's mfx_unweight): keep e iff w > 0 and
// u * wmax < w with u the first uniform of philox (ev0 + e, 0xffffffff, iteration).
int mfo_unweight(const double* w, int64_t n, double wmax, uint32_t seed, uint32_t iteration, uint64_t ev0,
                 uint8_t* keep) {
  for (int64_t e = 0; e < n; ++e) {
    u32x4 c;
    c.v[0] = (uint32_t)(ev0 + e);
    c.v[1] = (uint32_t)((ev0 + e) >> 32);
    c.v[2] = 0xffffffffu;
    c.v[3] = iteration;
    const u32x4 r = philox4x32_10(c, seed, 0x6d66u);
    const double u = u53(r.v[0], r.v[1]);
    keep[e] = (w[e] > 0.0 && u * wmax < w[e]) ? 1 : 0;
  }
  return 0;
}

int mfo_philox_raw(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  u32x4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  u32x4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
  return 0;
}

int mfo_vegas_map(const double* edges, int ndim, int nbins, const double* y, int64_t nev, double* x, double* jac,
                  int* bins) {
  VegasGrid g;
  g.ndim = ndim;
  g.nbins = nbins;
  g.edges.assign(edges, edges + (size_t)ndim * (nbins + 1));
  std::vector<int> b(ndim);
  for (int64_t e = 0; e < nev; ++e) {
    jac[e] = vegas_map(g, y + e * ndim, x + e * ndim, b.data());
    if (bins)
      for (int d = 0; d < ndim; ++d) bins[e * ndim + d] = b[d];
  }
  return 0;
}

int mfo_vegas_refine(double* edges, const double* hist, int ndim, int nbins, double alpha) {
  for (int d = 0; d < ndim; ++d) vegas_refine_dim(edges + (size_t)d * (nbins + 1), hist + (size_t)d * nbins, nbins, alpha);
  return 0;
}

// Oracle of one VEGAS iteration's event generation: same Philox stream, VEGAS map,
// RAMBO (mapping 0) or RAMBO on diet (mapping 1), cuts on legs cut_first_leg..n_out+1
// (pT > pt_min, |eta| < abs_eta_max, Delta R > dr_min), wps = w * jac * gev2pb / (2 s).
// mapping 2 (hadronic) follows ramboflow (phasespace.py:255-317): _get_x1x2 (:213-232) on
// the first two dimensions, RAMBO at sqrt(shat) on the rest, the flux 1/(2 shat), the cuts
// in the partonic frame (:503-512) and, with lab_frame, _boost_to_lab (:322-357).
int mfo_generate(int n_out, int ndim, int nbins, int mapping, double sqrts, const double* masses, double gev2pb,
                 double pt_min, double abs_eta_max, double dr_min, int cut_first_leg, uint32_t seed,
                 const double* edges, uint64_t ev0, int64_t nev, uint32_t iteration, double* all_ps, double* wps,
                 uint8_t* flag, double shat_min, int lab_frame, double* x12) {
  VegasGrid g;
  g.ndim = ndim;
  g.nbins = nbins;
  g.edges.assign(edges, edges + (size_t)ndim * (nbins + 1));
  std::vector<double> y(ndim), x(ndim), p(4 * n_out);
  std::vector<int> bins(ndim);
  bool massive = false;
  for (int i = 0; i < n_out; ++i) massive |= masses[i] != 0;
  const int next = n_out + 2;
  for (int64_t e = 0; e < nev; ++e) {
    philox_uniforms(ev0 + e, ndim, seed, iteration, y.data());
    double jac = vegas_map(g, y.data(), x.data(), bins.data());
    double lw, roots = sqrts, x1 = 1, x2 = 1, wtau = 1;
    if (mapping == 2) {
      const double s = sqrts * sqrts, taumin = shat_min / s, delta_tau = 1 - taumin;
      const double tau = x[0] * delta_tau + taumin;
      x1 = std::pow(tau, x[1]);
      x2 = tau / x1;
      wtau = delta_tau * -std::log(tau);
      roots = std::sqrt(x1 * x2 * s);
      lw = rambo_massless_event(x.data() + 2, n_out, roots, M_PI, p.data());
    } else if (mapping == 0) {
      lw = rambo_massless_event(x.data(), n_out, sqrts, M_PI, p.data());
    } else {
      lw = diet_massless(x.data(), n_out, sqrts, M_PI, p.data());
    }
    bool converged = true;
    if (massive) lw += massive_factor(p.data(), n_out, masses, roots, &converged);
    double w = wtau * std::exp(lw) / std::pow(2 * M_PI, 3 * n_out - 4);
    bool ok = std::isfinite(w) && converged;
    for (int i = cut_first_leg - 2; i < n_out; ++i) {
      const double* q = &p[4 * i];
      if (!(pt_of(q) > pt_min)) ok = false;
      if (abs_eta_max > 0 && !(std::fabs(eta_of(q)) < abs_eta_max)) ok = false;
      if (dr_min > 0)
        for (int j = i + 1; j < n_out; ++j)
          if (!(delta_r(q, &p[4 * j]) > dr_min)) ok = false;
    }
    double* o = all_ps + e * next * 4;
    o[0] = roots / 2; o[1] = 0; o[2] = 0; o[3] = roots / 2;
    o[4] = roots / 2; o[5] = 0; o[6] = 0; o[7] = -roots / 2;
    for (int k = 0; k < 4 * n_out; ++k) o[8 + k] = p[k];
    if (mapping == 2 && lab_frame) {
      // p_lab = p_com . B with B = [[cosh, 0, 0, -sinh], [0,1,0,0], [0,0,1,0], [-sinh, 0, 0, cosh]]
      const double eta = -0.5 * std::log(x1 / x2), ch = std::cosh(eta), sh = std::sinh(eta);
      for (int i = 0; i < next; ++i) {
        const double E = o[4 * i], pz = o[4 * i + 3];
        o[4 * i] = E * ch + pz * -sh;
        o[4 * i + 3] = E * -sh + pz * ch;
      }
    }
    if (x12) {
      x12[2 * e] = x1;
      x12[2 * e + 1] = x2;
    }
    wps[e] = ok ? w * jac * gev2pb / (2 * roots * roots) : 0.0;
    flag[e] = ok ? 1 : (converged ? 0 : 2);   // 2: the massive Newton iteration did not converge
  }
  return 0;
}

double mfo_delta_r(const double* a, const double* b) { return delta_r(a, b); }
double mfo_eta(const double* a) { return eta_of(a); }

}  // extern "C"
