// RAMBO phase space as per-lane FP64 device code.
//   python_package/madflow/phasespace.py:108-212 (Kleiss, Stirling, Ellis,
//   Comput. Phys. Commun. 40 (1986) 359): isotropic massless q_i from 4 uniforms each,
//   conformal boost + scale to sum p = (sqrts, 0), flat weight; massive particles by
//   rescaling the 3-momenta with xi solving sum_i sqrt(m_i^2 + xi^2 E_i^2) = sqrts.
// Unlike the reference's batch-global tf.while_loop (the batch stops when one event
// converges, :90-103), every lane runs its own Newton iteration to convergence and
// the energies are recomputed from the final xi.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfx {

constexpr int RAMBO_MAXN = 8;
constexpr double MFX_PI = 3.14159265358979323846;

struct RamboConst {
  int n;
  int massive;
  double masses[RAMBO_MAXN];
  double total_mass;
  double logw0;   // log(pi/2)(n-1) - 2 lgamma(n-1) - log(n-1)
  double inorm;   // 1 / (2 pi)^(3n-4)
};

// Newton iterations allowed for the massive rescaling factor xi
constexpr int NEWTON_MAX = 50;

__device__ __forceinline__ double massive_rescale(const RamboConst& rc, double sqrts, double (*p)[4], bool& converged);

// x: 4n uniforms (cos theta, phi, and two for the energy) per particle; returns weight.
// converged: false when the massive rescaling's Newton iteration did not reach
// |sum E - sqrts| <= 1e-14 sqrts in NEWTON_MAX steps (the event is then off shell).
__device__ __forceinline__ double rambo_event(const RamboConst& rc, const double* x, double sqrts,
                                              double (*p)[4], bool& converged) {
  const int n = rc.n;
  double Q[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const double costh = 2.0 * x[4 * i] - 1.0;
    const double sinth = sqrt(1.0 - costh * costh);
    double sphi, cphi;
    sincos(2.0 * MFX_PI * x[4 * i + 1], &sphi, &cphi);
    const double e = -log(x[4 * i + 2] * x[4 * i + 3]);
    p[i][0] = e;
    p[i][1] = e * sinth * sphi;
    p[i][2] = e * sinth * cphi;
    p[i][3] = e * costh;
    for (int k = 0; k < 4; ++k) Q[k] += p[i][k];
  }
  const double M = sqrt(Q[0] * Q[0] - (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]));
  const double iM = 1.0 / M;
  const double xs = sqrts * iM;
  const double b1 = -Q[1] * iM, b2 = -Q[2] * iM, b3 = -Q[3] * iM;
  const double gamma = Q[0] * iM;
  const double a = 1.0 / (1.0 + gamma);
  for (int i = 0; i < n; ++i) {
    const double bq = p[i][1] * b1 + p[i][2] * b2 + p[i][3] * b3;
    const double t = bq * a + p[i][0];
    p[i][0] = (p[i][0] * gamma + bq) * xs;
    p[i][1] = (p[i][1] + b1 * t) * xs;
    p[i][2] = (p[i][2] + b2 * t) * xs;
    p[i][3] = (p[i][3] + b3 * t) * xs;
  }
  double logw = rc.logw0 + (2 * n - 4) * log(sqrts);
  converged = true;
  if (rc.massive) logw += massive_rescale(rc, sqrts, p, converged);
  return exp(logw) * rc.inorm;
}

// Massive momenta from massless ones (phasespace.py:195-212): rescale the 3-momenta by
// xi with sum_i sqrt(m_i^2 + xi^2 E_i^2) = sqrts; returns the log of the weight factor
// xi^(2n-3) prod(|p_i|/E_i) / sum(|p_i|^2/E_i) * sqrts.  Needs sum m_i < sqrts (the
// callers check it); converged reports whether Newton met its tolerance.
__device__ __forceinline__ double massive_rescale(const RamboConst& rc, double sqrts, double (*p)[4], bool& converged) {
  const int n = rc.n;
  double logw = 0.0;
  converged = false;
  {
    // Newton for xi, starting above the root (f(xi0) >= 0 by Minkowski's inequality)
    double xi = sqrt(1.0 - (rc.total_mass / sqrts) * (rc.total_mass / sqrts));
    for (int it = 0; it < NEWTON_MAX; ++it) {
      const double x2 = xi * xi;
      double f0 = -sqrts, g0 = 0.0;
      for (int i = 0; i < n; ++i) {
        const double E2 = p[i][0] * p[i][0];
        const double nE = sqrt(rc.masses[i] * rc.masses[i] + E2 * x2);
        f0 += nE;
        g0 += E2 / nE;
      }
      if (fabs(f0) <= 1e-14 * sqrts) {
        converged = true;
        break;
      }
      xi -= f0 / (xi * g0);
    }
    const double x2 = xi * xi;
    double wt2 = 1.0, wt3 = 0.0;
    for (int i = 0; i < n; ++i) {
      const double v = p[i][0] * xi;
      const double nE = sqrt(rc.masses[i] * rc.masses[i] + p[i][0] * p[i][0] * x2);
      wt2 *= v / nE;
      wt3 += v * v / nE;
      p[i][0] = nE;
      p[i][1] *= xi;
      p[i][2] *= xi;
      p[i][3] *= xi;
    }
    logw += (2 * n - 3) * log(xi) + log(wt2 / wt3 * sqrts);
  }
  return logw;
}

// Solve F(u) = r for u in [0,1], F(u) = (k+1) u^k - k u^(k+1)  (k >= 1): the CDF of
// u = M_{i+1}^2 / M_i^2 in the massless n-body factorisation (Newton with bisection guard).
__device__ __host__ inline double diet_solve(double r, int k) {
  double lo = 0.0, hi = 1.0;
  double u = pow(r, 1.0 / k);  // leading behaviour near u = 0
  for (int it = 0; it < 60; ++it) {
    const double uk = pow(u, k - 1);
    const double F = (k + 1) * uk * u - k * uk * u * u - r;
    if (F > 0) hi = u; else lo = u;
    const double dF = (k + 1) * k * uk * (1.0 - u);
    double un = (dF > 0) ? u - F / dF : 0.5 * (lo + hi);
    if (!(un > lo && un < hi)) un = 0.5 * (lo + hi);
    if (fabs(un - u) <= 1e-15 * u) { u = un; break; }
    u = un;
  }
  return u;
}

// "RAMBO on diet" (S. Platzer, arXiv:1308.2922): n massless momenta from 3n-4 uniforms
// with the same flat weight as RAMBO, by a chain of two-body decays Q_i -> p_i + Q_{i+1}
// whose invariant masses are drawn from their exact marginal densities; massive
// particles are then obtained with RAMBO's xi rescaling.  x[0..n-3] fix the masses,
// x[n-2+2i], x[n-1+2i] the decay angles of step i.
__device__ __forceinline__ double diet_event(const RamboConst& rc, const double* x, double sqrts, double (*p)[4],
                                             bool& converged) {
  const int n = rc.n;
  double Q[4] = {sqrts, 0.0, 0.0, 0.0};
  double M = sqrts;
  for (int i = 0; i < n - 1; ++i) {
    double Mn = 0.0;
    if (i < n - 2) Mn = M * sqrt(diet_solve(x[i], n - 2 - i));
    const double q = (M * M - Mn * Mn) / (2.0 * M);
    const double costh = 2.0 * x[n - 2 + 2 * i] - 1.0;
    const double sinth = sqrt(fmax(0.0, 1.0 - costh * costh));
    double sphi, cphi;
    sincos(2.0 * MFX_PI * x[n - 1 + 2 * i], &sphi, &cphi);
    // momentum in the rest frame of Q, then boost with Q/M
    const double k0 = q, k1 = q * sinth * cphi, k2 = q * sinth * sphi, k3 = q * costh;
    const double b1 = Q[1] / M, b2 = Q[2] / M, b3 = Q[3] / M, g = Q[0] / M;
    const double bk = b1 * k1 + b2 * k2 + b3 * k3;
    const double c = bk / (g + 1.0) + k0;
    p[i][0] = g * k0 + bk;
    p[i][1] = k1 + b1 * c;
    p[i][2] = k2 + b2 * c;
    p[i][3] = k3 + b3 * c;
    for (int j = 0; j < 4; ++j) Q[j] -= p[i][j];
    M = Mn;
  }
  for (int j = 0; j < 4; ++j) p[n - 1][j] = Q[j];
  double logw = rc.logw0 + (2 * n - 4) * log(sqrts);
  converged = true;
  if (rc.massive) logw += massive_rescale(rc, sqrts, p, converged);
  return exp(logw) * rc.inorm;
}

// Pure boost of k by the velocity of Q (mass M); sign = -1 boosts into Q's rest frame.
__device__ __forceinline__ void boost_by(const double* k, const double* Q, double M, double sign, double* out) {
  const double b1 = sign * Q[1] / M, b2 = sign * Q[2] / M, b3 = sign * Q[3] / M, g = Q[0] / M;
  const double bk = b1 * k[1] + b2 * k[2] + b3 * k[3];
  const double c = bk / (g + 1.0) + k[0];
  out[0] = g * k[0] + bk;
  out[1] = k[1] + b1 * c;
  out[2] = k[2] + b2 * c;
  out[3] = k[3] + b3 * c;
}

// Inverse of diet_event: n momenta p (overwritten) in the rest frame of their sum
// (M, 0, 0, 0) -> the 3n-4 cube coordinates x that diet_event maps onto them, and the
// weight diet_event gives there.  Massive particles are first mapped back to the massless
// momenta q_i = p_i / xi, xi = sum |p_i| / M (the inverse of RAMBO's rescaling); the
// massive weight factor follows from the p_i alone.  The decay masses then give
// x[i] = F(M_{i+1}^2 / M_i^2) with F the CDF diet_solve inverts, and the direction of q_i
// in the rest frame of Q_i = q_i + ... + q_{n-1} gives the two angle coordinates.
__device__ __forceinline__ double diet_inverse(const RamboConst& rc, double M, double (*p)[4], double* x) {
  const int n = rc.n;
  double logw = rc.logw0 + (2 * n - 4) * log(M);
  if (rc.massive) {
    double sa = 0.0, prod = 1.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) {
      const double pa = sqrt(p[i][1] * p[i][1] + p[i][2] * p[i][2] + p[i][3] * p[i][3]);
      sa += pa;
      prod *= pa / p[i][0];
      s2 += pa * pa / p[i][0];
    }
    const double xi = sa / M;
    logw += (2 * n - 3) * log(xi) + log(prod / s2 * M);
    const double ixi = 1.0 / xi;
    for (int i = 0; i < n; ++i) {
      const double pa = sqrt(p[i][1] * p[i][1] + p[i][2] * p[i][2] + p[i][3] * p[i][3]);
      p[i][0] = pa * ixi;
      p[i][1] *= ixi;
      p[i][2] *= ixi;
      p[i][3] *= ixi;
    }
  }
  double Q[4] = {M, 0.0, 0.0, 0.0};
  double Mc = M;
  for (int i = 0; i < n - 1; ++i) {
    double Qn[4];
    for (int j = 0; j < 4; ++j) Qn[j] = Q[j] - p[i][j];
    double Mn = 0.0;
    if (i < n - 2) {
      Mn = sqrt(fmax(Qn[0] * Qn[0] - Qn[1] * Qn[1] - Qn[2] * Qn[2] - Qn[3] * Qn[3], 0.0));
      const double u = (Mn / Mc) * (Mn / Mc);
      const int k = n - 2 - i;
      double uk = 1.0;
      for (int j = 0; j < k; ++j) uk *= u;
      x[i] = (k + 1) * uk - k * uk * u;
    }
    double kr[4];
    boost_by(p[i], Q, Mc, -1.0, kr);
    const double ka = sqrt(kr[1] * kr[1] + kr[2] * kr[2] + kr[3] * kr[3]);
    x[n - 2 + 2 * i] = 0.5 * (kr[3] / ka + 1.0);
    double ph = atan2(kr[2], kr[1]);
    if (ph < 0.0) ph += 2.0 * MFX_PI;
    x[n - 1 + 2 * i] = ph / (2.0 * MFX_PI);
    for (int j = 0; j < 4; ++j) Q[j] = Qn[j];
    Mc = Mn;
  }
  return exp(logw) * rc.inorm;
}

}  // namespace mfx
