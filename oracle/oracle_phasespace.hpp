// ORACLE — test infrastructure only (see oracle_helas.hpp).
// CPU restatement of the phase-space and integration pieces of madflow's hot path:
//   python_package/madflow/phasespace.py:108-212  RAMBO (massless + massive Newton)
//   python_package/madflow/phasespace.py:405-478  cuts (pt/mt/mt2; eta and deltaR added)
//   VEGAS (external vegasflow, not in the reference tree): Lepage, J.Comput.Phys. 27
//   (1978) 192 — map, per-dimension bin histograms of (f J)^2, smoothing + damping
//   ((1-r)/ln(1/r))^alpha, rebinning, 1/sigma^2 iteration combination.
//   Philox4x32-10 (Salmon et al., SC'11) counter-based uniforms shared with the GPU.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>
#include <algorithm>

namespace oracle {

// ------------------------------------------------------------------ Philox4x32-10
struct u32x4 { uint32_t v[4]; };
inline u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * ctr.v[0];
    uint64_t p1 = (uint64_t)M1 * ctr.v[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u32x4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}
// 53-bit uniform in the open interval (0,1)
inline double u53(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}
// Uniforms for event ``ev`` (dimension d): counter (ev_lo, ev_hi, d/2, iteration), key (seed, 0x6d66)
inline void philox_uniforms(uint64_t ev, int ndim, uint32_t seed, uint32_t iteration, double* out) {
  for (int j = 0; 2 * j < ndim; ++j) {
    u32x4 c;
    c.v[0] = (uint32_t)ev;
    c.v[1] = (uint32_t)(ev >> 32);
    c.v[2] = (uint32_t)j;
    c.v[3] = iteration;
    u32x4 r = philox4x32_10(c, seed, 0x6d66u);
    out[2 * j] = u53(r.v[0], r.v[1]);
    if (2 * j + 1 < ndim) out[2 * j + 1] = u53(r.v[2], r.v[3]);
  }
}

// --------------------------------------------------------------------- constants
struct PSConsts {
  double pi;     // float_me(np.pi) in the reference (phasespace.py:16)
  double acc;    // float_me(1e-10) (phasespace.py:17)
  double gev2pb; // float_me(389379365.6) (phasespace.py:316)
};
inline PSConsts ps_consts(bool ref_compat) {
  PSConsts c;
  if (ref_compat) {
    c.pi = (double)(float)M_PI;
    c.acc = (double)(float)1e-10;
    c.gev2pb = (double)(float)389379365.6;
  } else {
    c.pi = M_PI;
    c.acc = 1e-10;
    c.gev2pb = 389379365.6;
  }
  return c;
}

// ------------------------------------------------------------------------ RAMBO
// _gen_unconstrained_momenta (phasespace.py:121-142)
inline void rambo_q(const double x[4], double pi, double q[4]) {
  double costh = 2.0 * x[0] - 1.0;
  double sinth = std::sqrt(1.0 - costh * costh);
  double phi = 2 * pi * x[1];
  double energy = -1.0 * std::log(x[2] * x[3]);
  q[0] = energy;
  q[1] = energy * sinth * std::sin(phi);
  q[2] = energy * sinth * std::cos(phi);
  q[3] = energy * costh;
}

// Massless RAMBO for one event (phasespace.py:171-193); returns log-weight before exp.
inline double rambo_massless_event(const double* xrand, int n, double sqrts, double pi, double* p /*n*4*/) {
  std::vector<double> q(4 * n);
  double sq[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    rambo_q(xrand + 4 * i, pi, &q[4 * i]);
    for (int k = 0; k < 4; ++k) sq[k] += q[4 * i + k];
  }
  double qmass = std::sqrt(sq[0] * sq[0] - (sq[1] * sq[1] + sq[2] * sq[2] + sq[3] * sq[3]));
  double x = sqrts / qmass;
  double b[4];
  for (int k = 0; k < 4; ++k) b[k] = -sq[k] / qmass;
  double gamma = -b[0];
  double a = 1.0 / (1.0 + gamma);
  for (int i = 0; i < n; ++i) {
    const double* qi = &q[4 * i];
    // _conformal_transformation (phasespace.py:108-118)
    double bq = qi[1] * b[1] + qi[2] * b[2] + qi[3] * b[3];
    double tmp = bq * a + qi[0];
    p[4 * i + 0] = (qi[0] * gamma + bq) * x;
    for (int k = 1; k < 4; ++k) p[4 * i + k] = (qi[k] + b[k] * tmp) * x;
  }
  double wt = std::log(pi / 2.0) * (n - 1);
  wt -= 2.0 * std::lgamma((double)(n - 1));
  wt -= std::log((double)(n - 1));
  wt += (2 * n - 4) * std::log(sqrts);
  return wt;
}

// Batched RAMBO (phasespace.py:145-212).
// ref_compat=true reproduces the batch-global Newton loop of _massive_xfactor (:71-105):
// the loop stops as soon as one event has f0 <= ACC (max 10 passes) and the returned
// energies lag the returned xfactor by one update.  ref_compat=false converges every
// event independently (|f0| <= 1e-14*sqrts, at most 50 Newton steps) and returns
// energies consistent with xfactor.
inline void rambo_batch(const double* xrand, int64_t nev, int xstride, int n, const double* sqrts, int sqrts_stride,
                        const double* masses, bool ref_compat, double* p_out, double* wt_out) {
  PSConsts C = ps_consts(ref_compat);
  bool massive = false;
  double total_mass = 0;
  if (masses) {
    for (int i = 0; i < n; ++i) {
      total_mass += masses[i];
      if (masses[i] != 0) massive = true;
    }
  }
  std::vector<double> logw(nev);
  for (int64_t e = 0; e < nev; ++e) {
    logw[e] = rambo_massless_event(xrand + e * xstride, n, sqrts[e * sqrts_stride], C.pi, p_out + e * n * 4);
  }
  if (!massive) {
    for (int64_t e = 0; e < nev; ++e) wt_out[e] = std::exp(logw[e]) / std::pow(2 * C.pi, 3 * n - 4);
    return;
  }
  std::vector<double> xf(nev), newE((size_t)nev * n);
  for (int64_t e = 0; e < nev; ++e) {
    double s = sqrts[e * sqrts_stride];
    xf[e] = std::sqrt(1 - (total_mass / s) * (total_mass / s));
    for (int i = 0; i < n; ++i) newE[e * n + i] = p_out[e * n * 4 + 4 * i];
  }
  if (ref_compat) {
    bool go = true;
    for (int it = 0; it < 10 && go; ++it) {
      bool all_gt = true;
      for (int64_t e = 0; e < nev; ++e) {
        double s = sqrts[e * sqrts_stride];
        double x2 = xf[e] * xf[e];
        double f0 = 0, g0 = 0;
        for (int i = 0; i < n; ++i) {
          double E = p_out[e * n * 4 + 4 * i];
          double nE = std::sqrt(masses[i] * masses[i] + E * E * x2);
          newE[e * n + i] = nE;
          f0 += nE;
          g0 += E * E / nE;
        }
        f0 -= s;
        double nx = xf[e] - f0 / (xf[e] * g0);
        bool gt = f0 > C.acc;
        if (gt) xf[e] = nx;
        else all_gt = false;
      }
      go = all_gt;
    }
  } else {
    for (int64_t e = 0; e < nev; ++e) {
      double s = sqrts[e * sqrts_stride];
      for (int it = 0; it < 50; ++it) {
        double x2 = xf[e] * xf[e];
        double f0 = 0, g0 = 0;
        for (int i = 0; i < n; ++i) {
          double E = p_out[e * n * 4 + 4 * i];
          double nE = std::sqrt(masses[i] * masses[i] + E * E * x2);
          f0 += nE;
          g0 += E * E / nE;
        }
        f0 -= s;
        if (std::fabs(f0) <= 1e-14 * s) break;
        xf[e] = xf[e] - f0 / (xf[e] * g0);
      }
      double x2 = xf[e] * xf[e];
      for (int i = 0; i < n; ++i) {
        double E = p_out[e * n * 4 + 4 * i];
        newE[e * n + i] = std::sqrt(masses[i] * masses[i] + E * E * x2);
      }
    }
  }
  for (int64_t e = 0; e < nev; ++e) {
    double s = sqrts[e * sqrts_stride];
    double wt2 = 1.0, wt3 = 0.0;
    for (int i = 0; i < n; ++i) {
      double* pi_ = p_out + e * n * 4 + 4 * i;
      double v = pi_[0] * xf[e];
      double nE = newE[e * n + i];
      wt2 *= v / nE;
      wt3 += v * v / nE;
      pi_[0] = nE;
      for (int k = 1; k < 4; ++k) pi_[k] *= xf[e];
    }
    double lw = logw[e] + (2 * n - 3) * std::log(xf[e]) + std::log(wt2 / wt3 * s);
    wt_out[e] = std::exp(lw) / std::pow(2 * C.pi, 3 * n - 4);
  }
}

// Massive rescaling of massless momenta p (n x 4) in place (phasespace.py:195-212),
// per-event Newton; returns the log weight factor.  *converged (if given) reports whether
// |sum E - sqrts| <= 1e-14 sqrts was reached within 50 steps.
inline double massive_factor(double* p, int n, const double* masses, double sqrts, bool* converged = nullptr) {
  double tm = 0;
  for (int i = 0; i < n; ++i) tm += masses[i];
  double xf = std::sqrt(1 - (tm / sqrts) * (tm / sqrts));
  if (converged) *converged = false;
  for (int it = 0; it < 50; ++it) {
    double x2 = xf * xf, f0 = 0, g0 = 0;
    for (int i = 0; i < n; ++i) {
      double E = p[4 * i];
      double nE = std::sqrt(masses[i] * masses[i] + E * E * x2);
      f0 += nE;
      g0 += E * E / nE;
    }
    f0 -= sqrts;
    if (std::fabs(f0) <= 1e-14 * sqrts) {
      if (converged) *converged = true;
      break;
    }
    xf = xf - f0 / (xf * g0);
  }
  double wt2 = 1.0, wt3 = 0.0, x2 = xf * xf;
  for (int i = 0; i < n; ++i) {
    double* q = p + 4 * i;
    double v = q[0] * xf;
    double nE = std::sqrt(masses[i] * masses[i] + q[0] * q[0] * x2);
    wt2 *= v / nE;
    wt3 += v * v / nE;
    q[0] = nE;
    for (int k = 1; k < 4; ++k) q[k] *= xf;
  }
  return (2 * n - 3) * std::log(xf) + std::log(wt2 / wt3 * sqrts);
}

// RAMBO on diet (S. Platzer, arXiv:1308.2922): 3n-4 uniforms -> n massless momenta with
// RAMBO's flat weight, via sequential two-body decays.  u = M_{i+1}^2/M_i^2 has CDF
// (k+1) u^k - k u^(k+1), k = n-2-i; solved here by plain bisection (60 halvings).
inline double diet_massless(const double* x, int n, double sqrts, double pi, double* p) {
  double Q[4] = {sqrts, 0, 0, 0};
  double M = sqrts;
  for (int i = 0; i < n - 1; ++i) {
    double Mn = 0.0;
    if (i < n - 2) {
      int k = n - 2 - i;
      double lo = 0, hi = 1;
      for (int it = 0; it < 60; ++it) {
        double u = 0.5 * (lo + hi);
        double F = (k + 1) * std::pow(u, k) - k * std::pow(u, k + 1);
        if (F > x[i]) hi = u; else lo = u;
      }
      Mn = M * std::sqrt(0.5 * (lo + hi));
    }
    double q = (M * M - Mn * Mn) / (2 * M);
    double ct = 2 * x[n - 2 + 2 * i] - 1, st = std::sqrt(std::max(0.0, 1 - ct * ct));
    double ph = 2 * pi * x[n - 1 + 2 * i];
    double k[4] = {q, q * st * std::cos(ph), q * st * std::sin(ph), q * ct};
    // boost from the rest frame of Q
    double b[3] = {Q[1] / M, Q[2] / M, Q[3] / M}, g = Q[0] / M;
    double bk = b[0] * k[1] + b[1] * k[2] + b[2] * k[3];
    double c = bk / (g + 1) + k[0];
    double* pi_ = p + 4 * i;
    pi_[0] = g * k[0] + bk;
    for (int j = 0; j < 3; ++j) pi_[1 + j] = k[1 + j] + b[j] * c;
    for (int j = 0; j < 4; ++j) Q[j] -= pi_[j];
    M = Mn;
  }
  for (int j = 0; j < 4; ++j) p[4 * (n - 1) + j] = Q[j];
  double wt = std::log(pi / 2.0) * (n - 1) - 2.0 * std::lgamma((double)(n - 1)) - std::log((double)(n - 1)) +
              (2 * n - 4) * std::log(sqrts);
  return wt;
}

// ------------------------------------------------------------------------- cuts
inline double pt_of(const double* p) { return std::sqrt(p[1] * p[1] + p[2] * p[2]); }
inline double eta_of(const double* p) {
  double pabs = std::sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  return 0.5 * std::log((pabs + p[3]) / (pabs - p[3]));
}
inline double phi_of(const double* p) { return std::atan2(p[2], p[1]); }
inline double delta_r(const double* a, const double* b) {
  double de = eta_of(a) - eta_of(b);
  double dp = std::fabs(phi_of(a) - phi_of(b));
  if (dp > M_PI) dp = 2 * M_PI - dp;
  return std::sqrt(de * de + dp * dp);
}

// ------------------------------------------------------------------------ VEGAS
struct VegasGrid {
  int ndim = 0, nbins = 50;
  std::vector<double> edges;  // ndim * (nbins+1)
  void init(int nd, int nb) {
    ndim = nd;
    nbins = nb;
    edges.assign((size_t)nd * (nb + 1), 0.0);
    for (int d = 0; d < nd; ++d)
      for (int b = 0; b <= nb; ++b) edges[(size_t)d * (nb + 1) + b] = (double)b / nb;
  }
};

// y in (0,1)^ndim -> x; returns the jacobian and the bin index per dimension
inline double vegas_map(const VegasGrid& g, const double* y, double* x, int* bins) {
  double jac = 1.0;
  for (int d = 0; d < g.ndim; ++d) {
    double xn = y[d] * g.nbins;
    int b = (int)xn;
    if (b >= g.nbins) b = g.nbins - 1;
    const double* e = &g.edges[(size_t)d * (g.nbins + 1)];
    double w = e[b + 1] - e[b];
    x[d] = e[b] + w * (xn - b);
    jac *= w * g.nbins;
    bins[d] = b;
  }
  return jac;
}

// Refine one dimension from its histogram of (f J)^2 (alpha damping, Lepage rebinning).
inline void vegas_refine_dim(double* edges, const double* hist, int nbins, double alpha) {
  std::vector<double> sm(nbins), reg(nbins);
  if (nbins == 1) return;
  sm[0] = (hist[0] + hist[1]) / 2.0;
  sm[nbins - 1] = (hist[nbins - 2] + hist[nbins - 1]) / 2.0;
  for (int i = 1; i < nbins - 1; ++i) sm[i] = (hist[i - 1] + hist[i] + hist[i + 1]) / 3.0;
  double sum = 0;
  for (int i = 0; i < nbins; ++i) sum += sm[i];
  if (!(sum > 0)) return;
  double rsum = 0;
  for (int i = 0; i < nbins; ++i) {
    double r = sm[i] / sum;
    reg[i] = (r > 0 && r < 1) ? std::pow((1.0 - r) / std::log(1.0 / r), alpha) : 0.0;
    rsum += reg[i];
  }
  if (!(rsum > 0)) return;
  double ave = rsum / nbins;
  std::vector<double> ne(nbins + 1);
  ne[0] = 0.0;
  ne[nbins] = 1.0;
  int k = -1;
  double acc = 0;
  for (int j = 1; j < nbins; ++j) {
    while (acc < ave) {
      ++k;
      acc += reg[k];
    }
    acc -= ave;
    double lo = edges[k], hi = edges[k + 1];
    ne[j] = hi - (hi - lo) * acc / reg[k];
  }
  for (int j = 0; j <= nbins; ++j) edges[j] = ne[j];
}

}  // namespace oracle
