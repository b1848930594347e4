// VEGAS integration on the device: the event loop of the reference's integrand
// (python_package/madflow/scripts/madflow_exec.py:419-472, utilities.py:82-86) with the
// VEGAS map / histogram / refinement that vegasflow provides (Lepage, J.Comput.Phys. 27
// (1978) 192), as HIP kernels.  Per iteration:
//   k_generate   Philox uniforms -> VEGAS map -> RAMBO or RAMBO-on-diet -> cuts ->
//                full momenta + phase-space weight + pass flag          (one event / lane)
//   k_scan       block counts -> exclusive offsets                       (one block)
//   k_compact    passing events -> dense all_ps for the smatrix kernel   (deterministic order)
//   (mfx_smatrix on the dense batch)
//   k_accumulate f = |M|^2 * wps per event; compensated FP64 sums of f, f^2; per-dimension
//                bin histograms of f^2 in a fixed order (bitwise reproducible)
//   k_reduce     fixed-order compensated reduction of the block partials
//   k_refine     grid refinement (smoothing, ((1-r)/ln(1/r))^alpha damping, rebinning)
#include "mfx_phasespace.h"
#include "mfx_rng.h"
#include "mfx_api_common.h"
#include "../../include/madflow_amd.h"
#include <math.h>
#include <string.h>

namespace mfx {

constexpr int GEN_BLOCK = 256;
constexpr int ACC_BLOCK = 256;
constexpr int ACC_GRID = 1024;      // fixed: the reduction order must not depend on nev
constexpr int MAXDIM = 32;
constexpr int MAXBINS = 64;
constexpr int NSCAL = 6;            // sum, comp, sum2, comp2, max, n_nonzero

struct GenConst {
  RamboConst rc;
  int ndim, nbins, mapping, nexternal;
  double sqrts, norm;              // norm = GeV^-2 -> pb conversion / (2 s)
  double gev2pb, shat_min;         // hadronic mapping: per-event flux, lower end of shat
  int lab_frame;
  double pt_min, abs_eta_max, dr_min;
  int cut_first, cut_last;         // outgoing-leg range the gluon cuts apply to
  uint32_t seed;
  int ncuts;                       // per-leg cuts of register_cut (outgoing leg index 0..)
  mfx_cut cuts[MFX_MAX_CUTS];
};

// --------------------------------------------------------------------- helpers
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// Neumaier-compensated accumulator
struct KSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    double t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(const KSum& o) {
    add(o.s);
    c += o.c;
  }
};

__device__ __forceinline__ void uniforms(const GenConst& g, uint64_t ev, uint32_t iteration, double* y) {
  for (int j = 0; 2 * j < g.ndim; ++j) {
    double u0, u1;
    philox_pair(ev, (uint32_t)j, g.seed, iteration, u0, u1);
    y[2 * j] = u0;
    if (2 * j + 1 < g.ndim) y[2 * j + 1] = u1;
  }
}

// A point of an external proposal: clamped into [0, 1] so that bin indices stay in
// range whatever the proposal returns; a non-finite coordinate maps to 0 and the event
// is rejected (returns false).
__device__ __forceinline__ bool external_point(const float* __restrict__ src, int ndim, double* y) {
  bool ok = true;
  for (int d = 0; d < ndim; ++d) {
    double v = (double)src[d];
    if (!isfinite(v)) {
      ok = false;
      v = 0.0;
    }
    y[d] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  }
  return ok;
}

__device__ __forceinline__ double vmap(const GenConst& g, const double* __restrict__ grid, const double* y, double* x,
                                       int* bins) {
  double jac = 1.0;
  const int nb = g.nbins;
  for (int d = 0; d < g.ndim; ++d) {
    const double xn = y[d] * nb;
    int b = (int)xn;
    if (b >= nb) b = nb - 1;
    const double lo = grid[d * (nb + 1) + b], hi = grid[d * (nb + 1) + b + 1];
    const double w = hi - lo;
    x[d] = lo + w * (xn - b);
    jac *= w * nb;
    if (bins) bins[d] = b;
  }
  return jac;
}

__device__ __forceinline__ double eta_of(const double* p) {
  const double pa = sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  return 0.5 * log((pa + p[3]) / (pa - p[3]));
}

__device__ __forceinline__ bool pass_cuts(const GenConst& g, double (*p)[4]) {
  for (int i = g.cut_first; i <= g.cut_last; ++i) {
    const double pt = sqrt(p[i][1] * p[i][1] + p[i][2] * p[i][2]);
    if (!(pt > g.pt_min)) return false;
    if (g.abs_eta_max > 0.0 && !(fabs(eta_of(p[i])) < g.abs_eta_max)) return false;
  }
  if (g.dr_min > 0.0) {
    for (int i = g.cut_first; i <= g.cut_last; ++i)
      for (int j = i + 1; j <= g.cut_last; ++j) {
        const double de = eta_of(p[i]) - eta_of(p[j]);
        double dp = fabs(atan2(p[i][2], p[i][1]) - atan2(p[j][2], p[j][1]));
        if (dp > MFX_PI) dp = 2.0 * MFX_PI - dp;
        if (!(de * de + dp * dp > g.dr_min * g.dr_min)) return false;
      }
  }
  // register_cut's predicates (phasespace.py:457-466): strict bounds on pt, mt or mt2
  for (int c = 0; c < g.ncuts; ++c) {
    const mfx_cut& k = g.cuts[c];
    const double* q = p[k.particle];
    const double pt2 = q[1] * q[1] + q[2] * q[2];
    double v;
    if (k.variable == 0) {
      v = sqrt(pt2);
    } else {
      const double mt2 = q[0] * q[0] - q[3] * q[3];      // m^2 + pt^2
      v = k.variable == 1 ? sqrt(fmax(mt2, 0.0)) : mt2;
    }
    if (k.use_min && !(v > k.min_val)) return false;
    if (k.use_max && !(v < k.max_val)) return false;
  }
  return true;
}

// ---------------------------------------------------------------- generation
__global__ void __launch_bounds__(GEN_BLOCK) k_generate(GenConst g, const double* __restrict__ grid, uint64_t ev0,
                                                        int64_t nev, uint32_t iteration, const float* __restrict__ y_in,
                                                        const double* __restrict__ w_in, const double* __restrict__ x_in,
                                                        double* __restrict__ all_ps,
                                                        double* __restrict__ wps, uint8_t* __restrict__ flag,
                                                        double* __restrict__ x12, int* __restrict__ block_count) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  bool ok = false;
  if (e < nev) {
    double y[MAXDIM], x[MAXDIM];
    bool y_ok = true;
    double jac = 1.0;
    if (x_in) {
      // the caller's FP64 points, used as they are (no VEGAS grid): a non-finite coordinate
      // rejects the event, the others are clamped to [0, 1]
      for (int d = 0; d < g.ndim; ++d) {
        double v = x_in[e * g.ndim + d];
        if (!isfinite(v)) {
          y_ok = false;
          v = 0.0;
        }
        x[d] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
      }
    } else {
      if (y_in) {
        y_ok = external_point(y_in + e * g.ndim, g.ndim, y);
      } else {
        uniforms(g, ev0 + e, iteration, y);
      }
      jac = vmap(g, grid, y, x, nullptr);
    }
    double p[RAMBO_MAXN + 2][4];           // incoming partons, then the outgoing legs
    double w, norm = g.norm, x1 = 1.0, x2 = 1.0, eh = 0.5 * g.sqrts;
    bool converged = true;
    if (g.mapping == 2) {
      // ramboflow (phasespace.py:213-232,255-317): tau = x0 (1 - tau_min) + tau_min,
      // x1 = tau^x1', x2 = tau / x1, weight (1 - tau_min) (-ln tau); RAMBO at
      // sqrt(shat) = sqrt(x1 x2 s) on the other 4 n_out dimensions; flux 1 / (2 shat)
      const double s = g.sqrts * g.sqrts, taumin = g.shat_min / s, dtau = 1.0 - taumin;
      const double tau = x[0] * dtau + taumin;
      x1 = pow(tau, x[1]);
      x2 = tau / x1;
      const double shat = x1 * x2 * s;
      w = dtau * -log(tau) * rambo_event(g.rc, x + 2, sqrt(shat), p + 2, converged);
      norm = g.gev2pb / (2.0 * shat);
      eh = 0.5 * sqrt(shat);
    } else {
      w = g.mapping == 0 ? rambo_event(g.rc, x, g.sqrts, p + 2, converged)
                         : diet_event(g.rc, x, g.sqrts, p + 2, converged);
    }
    p[0][0] = eh; p[0][1] = 0.0; p[0][2] = 0.0; p[0][3] = eh;
    p[1][0] = eh; p[1][1] = 0.0; p[1][2] = 0.0; p[1][3] = -eh;
    ok = y_ok && converged && pass_cuts(g, p + 2) && isfinite(w);
    if (g.mapping == 2 && g.lab_frame) {
      // boost along z to the lab frame, rapidity -eta = (1/2) ln(x1 / x2) (phasespace.py:322-357)
      const double eta = -0.5 * log(x1 / x2), ch = cosh(eta), sh = sinh(eta);
      for (int i = 0; i < g.nexternal; ++i) {
        const double E = p[i][0], pz = p[i][3];
        p[i][0] = ch * E - sh * pz;
        p[i][3] = ch * pz - sh * E;
      }
    }
    double* out = all_ps + e * (g.nexternal * 4);
    for (int i = 0; i < g.nexternal; ++i)
      for (int k = 0; k < 4; ++k) out[4 * i + k] = p[i][k];
    if (x12) {
      x12[2 * e] = x1;
      x12[2 * e + 1] = x2;
    }
    const double wext = w_in ? w_in[e] : 1.0;
    ok = ok && isfinite(wext);
    wps[e] = ok ? w * jac * norm * wext : 0.0;
    // 1 passes, 0 cut, 2 the massive Newton iteration did not converge (off shell)
    flag[e] = ok ? 1 : (converged ? 0 : 2);
  }
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long bal = __ballot(ok);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, __popcll(bal));
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = cnt;
}

// ---------------------------------------------------------------- multi-channel
// Multi-channel phase space (Kleiss and Pittau, Comput. Phys. Commun. 83 (1994) 141;
// one VEGAS grid per channel as in T. Ohl, Comput. Phys. Commun. 120 (1999) 13).  The
// density of a point is g(p) = sum_j alpha_j g_j(p), g_j = 1 / (VEGAS jacobian of grid j
// at x_j(p) x dPhi/dx of channel j's map at p), where x_j(p) is the inverse of channel j's
// map; the weight is 1 / g(p).  Maps (outgoing legs 0 = t, 1 = t~, 2.. = gluons):
//   0 flat:  RAMBO on diet over every outgoing leg;
//   1 t-channel of leg a: pT ~ 1/pT^2 on [lo, hi = E_max], rapidity flat in the kinematic
//     range |y| <= acosh(E_max / pT), azimuth flat; the other legs by diet in the rest
//     frame of P - k_a  (Phi_n = int d^3k / ((2pi)^3 2E) Phi_{n-1}(P - k));
//   2 s-channel of legs (a, b): m_ab^2 ~ 1/m^2 on [lo, hi], isotropic decay in the pair's
//     rest frame; the other legs and the pair by diet in the c.m. frame
//     (Phi_n = int dm^2 / (2pi) Phi_{n-1}(P; ..., G(m)) Phi_2(G), Phi_2 = 1 / (8 pi)).
constexpr int MC_MAXCH = 8;

struct McChannel {
  int type, a, b, nsub;
  int legs[RAMBO_MAXN];   // outgoing leg of each diet slot; the pair G is the last slot
  double lo, hi;
  RamboConst sub;         // diet constants of the sub-system (s-channel: G's mass per event)
};

struct McConst {
  int nch;
  double scale;           // weight factor alpha_c N / n_c of the stratified estimator
  double alpha[MC_MAXCH];
  McChannel ch[MC_MAXCH];
};

__device__ __forceinline__ double grid_jacobian(const double* __restrict__ grid, int ndim, int nb, const double* x) {
  double jac = 1.0;
  for (int d = 0; d < ndim; ++d) {
    const double* e = grid + d * (nb + 1);
    int lo = 0, hi = nb;            // e[lo] <= x < e[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (e[mid] <= x[d]) lo = mid; else hi = mid;
    }
    jac *= (e[lo + 1] - e[lo]) * nb;
  }
  return jac;
}

// Upper end of a channel's range at the partonic energy sqrts: the range given, capped by
// the kinematic limit there (t-channel: the largest energy of the massless leg,
// (s - T^2) / (2 sqrt(s)); s-channel: (sqrt(s) - T)^2, with T the other legs' masses).  At
// fixed energy (mapping 1) the range given is used as it is.
__device__ __forceinline__ double mc_hi(const McChannel& ch, int mapping, double sqrts) {
  if (mapping != 3 || ch.type == 0) return ch.hi;
  const double T = ch.sub.total_mass;
  const double lim = ch.type == 1 ? (sqrts * sqrts - T * T) / (2.0 * sqrts) : (sqrts - T) * (sqrts - T);
  return fmin(ch.hi, lim);
}

// channel map: cube point x -> outgoing momenta p (c.m. frame); returns dPhi/dx.  hi is the
// upper end of the channel's range at this energy (mc_hi); conv = false when the range is
// empty there
__device__ double mc_forward(const McChannel& ch, double hi, double sqrts, const double* x, double (*p)[4],
                             bool& conv) {
  conv = true;
  if (ch.type == 0) return diet_event(ch.sub, x, sqrts, p, conv);
  if (!(hi > ch.lo)) {
    conv = false;
    return 0.0;
  }
  double q[RAMBO_MAXN][4];
  if (ch.type == 1) {
    const double A = 1.0 / ch.lo, B = 1.0 / hi;
    const double pt = 1.0 / (A - x[0] * (A - B));
    const double ymax = acosh(fmax(hi / pt, 1.0));
    const double y = (2.0 * x[1] - 1.0) * ymax;
    double sp, cp;
    sincos(2.0 * MFX_PI * x[2], &sp, &cp);
    const double k[4] = {pt * cosh(y), pt * cp, pt * sp, pt * sinh(y)};
    const double X[4] = {sqrts - k[0], -k[1], -k[2], -k[3]};
    const double MX = sqrt(fmax(X[0] * X[0] - X[1] * X[1] - X[2] * X[2] - X[3] * X[3], 0.0));
    if (!(MX > ch.sub.total_mass)) {
      conv = false;
      return 0.0;
    }
    const double w = diet_event(ch.sub, x + 3, MX, q, conv);
    for (int j = 0; j < ch.nsub; ++j) boost_by(q[j], X, MX, 1.0, p[ch.legs[j]]);
    for (int c = 0; c < 4; ++c) p[ch.a][c] = k[c];
    return pt * pt * pt * (A - B) * (2.0 * ymax) / (8.0 * MFX_PI * MFX_PI) * w;
  }
  const double L = log(hi / ch.lo);
  const double m2 = ch.lo * exp(x[0] * L), m = sqrt(m2);
  RamboConst rc = ch.sub;
  rc.masses[ch.nsub - 1] = m;
  rc.total_mass += m;
  rc.massive = 1;
  if (!(sqrts > rc.total_mass)) {
    conv = false;
    return 0.0;
  }
  const double w = diet_event(rc, x + 3, sqrts, q, conv);
  const double* G = q[ch.nsub - 1];
  const double ct = 2.0 * x[1] - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct));
  double sp, cp;
  sincos(2.0 * MFX_PI * x[2], &sp, &cp);
  const double h = 0.5 * m;
  const double ka[4] = {h, h * st * cp, h * st * sp, h * ct}, kb[4] = {h, -h * st * cp, -h * st * sp, -h * ct};
  for (int j = 0; j < ch.nsub - 1; ++j)
    for (int c = 0; c < 4; ++c) p[ch.legs[j]][c] = q[j][c];
  boost_by(ka, G, m, 1.0, p[ch.a]);
  boost_by(kb, G, m, 1.0, p[ch.b]);
  return w * m2 * L / (16.0 * MFX_PI * MFX_PI);
}

// inverse of mc_forward: outgoing momenta (c.m. frame) -> x and dPhi/dx; ok = false
// outside the channel's support (its density is zero there)
__device__ double mc_inverse(const McChannel& ch, double hi, double sqrts, const double (*p)[4], double* x, bool& ok) {
  double q[RAMBO_MAXN][4];
  ok = true;
  if (ch.type == 0) {
    for (int j = 0; j < ch.nsub; ++j)
      for (int c = 0; c < 4; ++c) q[j][c] = p[j][c];
    return diet_inverse(ch.sub, sqrts, q, x);
  }
  if (!(hi > ch.lo)) {
    ok = false;
    return 0.0;
  }
  if (ch.type == 1) {
    const double* k = p[ch.a];
    const double pt = sqrt(k[1] * k[1] + k[2] * k[2]);
    ok = pt >= ch.lo && pt <= hi;
    const double ptc = fmin(fmax(pt, ch.lo), hi);
    const double A = 1.0 / ch.lo, B = 1.0 / hi;
    const double ymax = acosh(fmax(hi / ptc, 1.0));
    const double y = asinh(k[3] / ptc);
    ok = ok && fabs(y) <= ymax;
    x[0] = (A - 1.0 / ptc) / (A - B);
    x[1] = 0.5 * (ymax > 0.0 ? y / ymax : 0.0) + 0.5;
    double ph = atan2(k[2], k[1]);
    if (ph < 0.0) ph += 2.0 * MFX_PI;
    x[2] = ph / (2.0 * MFX_PI);
    const double X[4] = {sqrts - k[0], -k[1], -k[2], -k[3]};
    const double MX = sqrt(fmax(X[0] * X[0] - X[1] * X[1] - X[2] * X[2] - X[3] * X[3], 0.0));
    ok = ok && MX > ch.sub.total_mass;
    if (!ok) return 0.0;
    for (int j = 0; j < ch.nsub; ++j) boost_by(p[ch.legs[j]], X, MX, -1.0, q[j]);
    const double w = diet_inverse(ch.sub, MX, q, x + 3);
    return ptc * ptc * ptc * (A - B) * (2.0 * ymax) / (8.0 * MFX_PI * MFX_PI) * w;
  }
  double G[4];
  for (int c = 0; c < 4; ++c) G[c] = p[ch.a][c] + p[ch.b][c];
  const double m2 = G[0] * G[0] - G[1] * G[1] - G[2] * G[2] - G[3] * G[3];
  ok = m2 >= ch.lo && m2 <= hi;
  if (!ok) return 0.0;
  const double L = log(hi / ch.lo), m = sqrt(m2);
  x[0] = log(m2 / ch.lo) / L;
  double kr[4];
  boost_by(p[ch.a], G, m, -1.0, kr);
  const double ka = sqrt(kr[1] * kr[1] + kr[2] * kr[2] + kr[3] * kr[3]);
  x[1] = 0.5 * (kr[3] / ka + 1.0);
  double ph = atan2(kr[2], kr[1]);
  if (ph < 0.0) ph += 2.0 * MFX_PI;
  x[2] = ph / (2.0 * MFX_PI);
  for (int j = 0; j < ch.nsub - 1; ++j)
    for (int c = 0; c < 4; ++c) q[j][c] = p[ch.legs[j]][c];
  for (int c = 0; c < 4; ++c) q[ch.nsub - 1][c] = G[c];
  RamboConst rc = ch.sub;
  rc.masses[ch.nsub - 1] = m;
  rc.total_mass += m;
  rc.massive = 1;
  const double w = diet_inverse(rc, sqrts, q, x + 3);
  return w * m2 * L / (16.0 * MFX_PI * MFX_PI);
}

// VEGAS jacobian of a grid at the mapped point x, and (y != nullptr) the cube point y that
// the grid maps onto x
__device__ __forceinline__ double grid_inverse(const double* __restrict__ grid, int ndim, int nb, const double* x,
                                               double* y) {
  double jac = 1.0;
  for (int d = 0; d < ndim; ++d) {
    const double* e = grid + d * (nb + 1);
    int lo = 0, hi = nb;            // e[lo] <= x < e[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (e[mid] <= x[d]) lo = mid; else hi = mid;
    }
    const double w = e[lo + 1] - e[lo];
    jac *= w * nb;
    if (y) {
      const double t = w > 0.0 ? (x[d] - e[lo]) / w : 0.0;
      const double v = (lo + fmin(fmax(t, 0.0), 1.0)) / nb;
      y[d] = v;
    }
  }
  return jac;
}

// Points of channel c: VEGAS map of grid c, channel c's map, the densities of every
// channel at the point (ratios[e][j] = g_j / g) and the weight scale / g; then cuts, as
// k_generate (fixed sqrt(s), c.m. frame).
// External points (y_in, with w_in = 1 / their density on channel c's cube): the weight is
// multiplied by w_in, so that a proposal on the cube other than the uniform one leaves the
// estimate unbiased.
// Learned mode (y0 != nullptr; channel 0 is the flat channel): the points of channel c
// come from y_in when it is given (the learned channel's samples on channel 0's cube),
// and the kernel leaves the division by g to mfx_mc_combine, once the learned density at
// each point's channel-0 cube point y0 is known: it writes y0, the raw densities g_j in
// ratios (row stride nch + 1), gdens[e] = (sum_j alpha_j g_j, g_0) and wps = scale x norm.
__global__ void __launch_bounds__(GEN_BLOCK) k_mc_generate(GenConst g, McConst mc, int c,
                                                           const double* __restrict__ grids, uint64_t ev0, int64_t nev,
                                                           uint32_t iteration, const float* __restrict__ y_in,
                                                           const double* __restrict__ w_in,
                                                           double* __restrict__ all_ps, double* __restrict__ wps,
                                                           uint8_t* __restrict__ flag, double* __restrict__ ratios,
                                                           float* __restrict__ y0_out, double* __restrict__ gdens,
                                                           double* __restrict__ x12, int* __restrict__ block_count) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  bool ok = false;
  if (e < nev) {
    const int D = g.ndim, NB = g.nbins;
    const bool learned = y0_out != nullptr;
    const int rs = learned ? mc.nch + 1 : mc.nch;
    double y[MAXDIM], x[MAXDIM], y0[MAXDIM];
    bool y_ok = true;
    if (y_in) y_ok = external_point(y_in + e * D, D, y);
    else uniforms(g, ev0 + e, iteration, y);
    const double jac = vmap(g, grids + (size_t)c * D * (NB + 1), y, x, nullptr);
    // hadronic mode (mapping 3): the first two coordinates give tau = x1 x2 and x1 as in
    // the ramboflow mapping (weight (1 - tau_min)(-ln tau), the same for every channel), the
    // channels act at sqrt(shat) on the others
    double rs_hat = g.sqrts, norm = g.norm, wtau = 1.0, x1 = 1.0, x2 = 1.0;
    const int off = g.mapping == 3 ? 2 : 0;
    if (g.mapping == 3) {
      const double s = g.sqrts * g.sqrts, taumin = g.shat_min / s, dtau = 1.0 - taumin;
      const double tau = x[0] * dtau + taumin;
      x1 = pow(tau, x[1]);
      x2 = tau / x1;
      rs_hat = sqrt(tau * s);
      wtau = dtau * -log(tau);
      norm = g.gev2pb / (2.0 * tau * s);
    }
    double p[RAMBO_MAXN + 2][4];
    bool conv = true;
    const double w = mc_forward(mc.ch[c], mc_hi(mc.ch[c], g.mapping, rs_hat), rs_hat, x + off, p + 2, conv);
    double gj[MC_MAXCH], gsum = 0.0;
    for (int j = 0; j < mc.nch; ++j) {
      double v = 0.0;
      if (j == c) {
        v = 1.0 / (jac * w);
        if (learned && j == 0)
          for (int d = 0; d < D; ++d) y0[d] = y[d];
      } else if (conv) {
        double xj[MAXDIM];
        bool in = false;
        if (off) {
          xj[0] = x[0];
          xj[1] = x[1];
        }
        const double wj = mc_inverse(mc.ch[j], mc_hi(mc.ch[j], g.mapping, rs_hat), rs_hat, p + 2, xj + off, in);
        const double J = grid_inverse(grids + (size_t)j * D * (NB + 1), D, NB, xj, (learned && j == 0) ? y0 : nullptr);
        if (in && wj > 0.0) v = 1.0 / (J * wj);
      } else if (learned && j == 0) {
        for (int d = 0; d < D; ++d) y0[d] = 0.5;
      }
      if (!isfinite(v)) v = 0.0;
      gj[j] = v;
      gsum += mc.alpha[j] * v;
    }
    const double eh = 0.5 * rs_hat;
    p[0][0] = eh; p[0][1] = 0.0; p[0][2] = 0.0; p[0][3] = eh;
    p[1][0] = eh; p[1][1] = 0.0; p[1][2] = 0.0; p[1][3] = -eh;
    const double wt = learned ? mc.scale : mc.scale * wtau / gsum;
    ok = y_ok && conv && w > 0.0 && (learned ? gj[0] > 0.0 : gsum > 0.0) && isfinite(wt) && pass_cuts(g, p + 2);
    if (g.mapping == 3 && g.lab_frame) {
      // boost along z to the lab frame, rapidity -eta = (1/2) ln(x1 / x2), after the cuts
      const double eta = -0.5 * log(x1 / x2), ch = cosh(eta), sh = sinh(eta);
      for (int i = 0; i < g.nexternal; ++i) {
        const double E = p[i][0], pz = p[i][3];
        p[i][0] = ch * E - sh * pz;
        p[i][3] = ch * pz - sh * E;
      }
    }
    double* out = all_ps + e * (g.nexternal * 4);
    for (int i = 0; i < g.nexternal; ++i)
      for (int k = 0; k < 4; ++k) out[4 * i + k] = p[i][k];
    if (x12) {
      x12[2 * e] = x1;
      x12[2 * e + 1] = x2;
    }
    const double wx = w_in ? w_in[e] : 1.0;    // 1 / (density of y on the cube) for external points
    ok = ok && isfinite(wx) && wx > 0.0;
    wps[e] = ok ? wt * norm * wx : 0.0;
    flag[e] = ok ? 1 : (conv ? 0 : 2);
    if (learned) {
      for (int j = 0; j < mc.nch; ++j) ratios[e * rs + j] = gj[j];
      ratios[e * rs + mc.nch] = 0.0;
      for (int d = 0; d < D; ++d) y0_out[e * D + d] = (float)fmin(fmax(y0[d], 0.0), 1.0);
      gdens[2 * e] = gsum;
      gdens[2 * e + 1] = gj[0];
    } else {
      for (int j = 0; j < mc.nch; ++j) ratios[e * rs + j] = gsum > 0.0 ? gj[j] / gsum : 0.0;
    }
  }
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long bal = __ballot(ok);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, __popcll(bal));
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = cnt;
}

// The learned channel's share of the density: g = gdens0 + alpha_learned q(y0) gdens1 with
// q = exp(lq) the learned density on channel 0's cube; wps /= g, the raw densities in
// ratios become g_j / g, and the last column q g_0 / g.
__global__ void __launch_bounds__(GEN_BLOCK) k_mc_combine(int64_t n, int nch, double alpha_learned,
                                                          const float* __restrict__ lq, const double* __restrict__ gdens,
                                                          double* __restrict__ wps, double* __restrict__ ratios) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  if (e >= n) return;
  const int rs = nch + 1;
  const double qg0 = exp((double)lq[e]) * gdens[2 * e + 1];
  const double gt = gdens[2 * e] + alpha_learned * qg0;
  const double inv = gt > 0.0 && isfinite(gt) ? 1.0 / gt : 0.0;
  const double w = wps[e] * inv;
  wps[e] = isfinite(w) ? w : 0.0;
  for (int j = 0; j < nch; ++j) ratios[e * rs + j] *= inv;
  ratios[e * rs + nch] = isfinite(qg0 * inv) ? qg0 * inv : 0.0;
}

// sum over events of f^2 r_j, f = me[slot] * wps: the channel-weight moments
// W_j = <(f/g)^2 g_j/g> of the Kleiss-Pittau update (per-block compensated partials)
__global__ void __launch_bounds__(ACC_BLOCK) k_mc_moments(int64_t nev, int nch, const int* __restrict__ slot,
                                                          const double* __restrict__ me, const double* __restrict__ wps,
                                                          const double* __restrict__ ratios,
                                                          double* __restrict__ partials) {
  KSum acc[MC_MAXCH];
  for (int64_t e = (int64_t)blockIdx.x * ACC_BLOCK + threadIdx.x; e < nev; e += (int64_t)gridDim.x * ACC_BLOCK) {
    const int s = slot[e];
    double f = s >= 0 ? me[s] * wps[e] : 0.0;
    if (!isfinite(f)) f = 0.0;
    for (int j = 0; j < nch; ++j) acc[j].add(f * f * ratios[e * nch + j]);
  }
  __shared__ double red[ACC_BLOCK][2];
  for (int j = 0; j < nch; ++j) {
    red[threadIdx.x][0] = acc[j].s;
    red[threadIdx.x][1] = acc[j].c;
    __syncthreads();
    for (int off = ACC_BLOCK / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
        KSum a{red[threadIdx.x][0], red[threadIdx.x][1]};
        a.merge(KSum{red[threadIdx.x + off][0], red[threadIdx.x + off][1]});
        red[threadIdx.x][0] = a.s;
        red[threadIdx.x][1] = a.c;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      partials[((size_t)blockIdx.x * nch + j) * 2] = red[0][0];
      partials[((size_t)blockIdx.x * nch + j) * 2 + 1] = red[0][1];
    }
    __syncthreads();
  }
}

__global__ void k_mc_reduce(int nparts, int nch, const double* __restrict__ partials, double* __restrict__ result) {
  const int j = threadIdx.x;
  if (j >= nch) return;
  KSum a;
  for (int b = 0; b < nparts; ++b) a.merge(KSum{partials[((size_t)b * nch + j) * 2], partials[((size_t)b * nch + j) * 2 + 1]});
  result[j] = a.s + a.c;
}

// ---------------------------------------------------------------- unweighting
// Hit-or-miss unweighting (the reference hands its weighted events to MadGraph's
// lhe_parser unweight, lhe_writer.py:319-356): event e is kept with probability
// min(1, w_e / wmax), drawn from the Philox stream (ev0 + e, UNW_STREAM, iteration) of the
// run's seed; kept events keep their order.  Events with w > wmax are always kept and
// carry weight w / wmax (the caller reads it from w), all others weight 1.
constexpr uint32_t UNW_STREAM = 0xffffffffu;
__global__ void __launch_bounds__(GEN_BLOCK) k_unweight_flag(const double* __restrict__ w, int64_t n, double wmax,
                                                             uint64_t ev0, uint32_t seed, uint32_t iteration,
                                                             uint8_t* __restrict__ keep, int* __restrict__ block_count) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  bool ok = false;
  if (e < n) {
    double u0, u1;
    philox_pair(ev0 + e, UNW_STREAM, seed, iteration, u0, u1);
    const double we = w[e];
    ok = we > 0.0 && u0 * wmax < we;
    keep[e] = ok ? 1 : 0;
  }
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long bal = __ballot(ok);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, __popcll(bal));
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = cnt;
}

__global__ void __launch_bounds__(GEN_BLOCK) k_unweight_compact(const uint8_t* __restrict__ keep, int64_t n,
                                                                const int* __restrict__ offsets,
                                                                int* __restrict__ index_out) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  const bool ok = e < n && keep[e];
  __shared__ int wave_base[GEN_BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(ok);
  if (lane == 0) wave_base[wave] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int v = 0; v < GEN_BLOCK / 64; ++v) {
      const int c = wave_base[v];
      wave_base[v] = acc;
      acc += c;
    }
  }
  __syncthreads();
  if (ok) index_out[offsets[blockIdx.x] + wave_base[wave] + __popcll(bal & ((1ull << lane) - 1ull))] = (int)e;
}

// exclusive scan of the block counts, one block; total in offsets[nblk]
__global__ void k_scan(const int* __restrict__ counts, int nblk, int* __restrict__ offsets) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nblk + 1023) / 1024;
  const int lo = t * per, hi = min(nblk, lo + per);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += counts[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int i = 0; i < 1024; ++i) {
      const int v = part[i];
      part[i] = acc;
      acc += v;
    }
    offsets[nblk] = acc;
  }
  __syncthreads();
  int acc = part[t];
  for (int i = lo; i < hi; ++i) {
    offsets[i] = acc;
    acc += counts[i];
  }
}

__global__ void __launch_bounds__(GEN_BLOCK) k_compact(int nexternal, int64_t nev, const double* __restrict__ all_ps,
                                                       const uint8_t* __restrict__ flag, const int* __restrict__ offsets,
                                                       double* __restrict__ dense, int* __restrict__ slot) {
  const int64_t e = (int64_t)blockIdx.x * GEN_BLOCK + threadIdx.x;
  const bool ok = e < nev && flag[e] == 1;
  __shared__ int wave_base[GEN_BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(ok);
  if (lane == 0) wave_base[wave] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < GEN_BLOCK / 64; ++w) {
      const int v = wave_base[w];
      wave_base[w] = acc;
      acc += v;
    }
  }
  __syncthreads();
  // The passing events of a wave land on consecutive ranks, so the wave copies their rows
  // together: lane l moves doubles l, l + 64, ... of the wave's dense range, which makes the
  // stores contiguous and the loads contiguous within each source row.
  __shared__ int src_lane[GEN_BLOCK];
  const int r = __popcll(bal & ((1ull << lane) - 1ull));
  const int64_t rank0 = (int64_t)offsets[blockIdx.x] + wave_base[wave];
  if (e < nev) slot[e] = ok ? (int)(rank0 + r) : -1;
  if (ok) src_lane[wave * 64 + r] = lane;
  __syncthreads();
  const int W = nexternal * 4, cnt = __popcll(bal);
  const int64_t e0 = (int64_t)blockIdx.x * GEN_BLOCK + wave * 64;
  const double* src = all_ps + e0 * W;
  double* dst = dense + rank0 * W;
  for (int j = lane; j < cnt * W; j += 64) {
    const int row = j / W;
    dst[j] = src[(int64_t)src_lane[wave * 64 + row] * W + (j - row * W)];
  }
}

// ---------------------------------------------------------------- accumulation
// partial layout per block: NSCAL scalars then ndim*nbins histogram (plain sums) and the
// same number of compensation terms.
__global__ void __launch_bounds__(ACC_BLOCK) k_accumulate(GenConst g, uint64_t ev0, int64_t nev, uint32_t iteration,
                                                          const float* __restrict__ y_in, const int* __restrict__ slot,
                                                          const double* __restrict__ me,
                                                          const double* __restrict__ wps, const double* __restrict__ fext,
                                                          double* __restrict__ fout, double* __restrict__ partials) {
  extern __shared__ double sh[];
  const int D = g.ndim, NB = g.nbins, HB = D * NB;
  const int nwave = ACC_BLOCK / 64;
  double* hist = sh;                                  // [nwave][HB]
  double* wv = sh + nwave * HB;                       // [nwave][64]
  uint8_t* wb = (uint8_t*)(wv + nwave * 64);          // [nwave][64][D] bin indices
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < nwave * HB; i += ACC_BLOCK) hist[i] = 0.0;
  __syncthreads();
  KSum s1, s2;
  double fmx = 0.0, nz = 0.0;
  const int64_t stride = (int64_t)gridDim.x * ACC_BLOCK;
  // every wave walks whole 64-event tiles so that lane order fixes the summation order
  for (int64_t base = (int64_t)blockIdx.x * ACC_BLOCK + wave * 64; base < nev; base += stride) {
    const int64_t e = base + lane;
    double f = 0.0;
    int bins[MAXDIM];
    if (e < nev) {
      if (fext) f = fext[e];
      else {
        const int s = slot[e];
        f = s >= 0 ? me[s] * wps[e] : 0.0;
      }
      if (!isfinite(f)) f = 0.0;          // a NaN/inf weight must not poison the grid
      double y[MAXDIM];
      if (y_in) {
        external_point(y_in + e * D, D, y);
      } else {
        uniforms(g, ev0 + e, iteration, y);
      }
      for (int d = 0; d < D; ++d) {
        int b = (int)(y[d] * NB);
        bins[d] = b >= NB ? NB - 1 : b;
      }
      if (fout) fout[e] = f;
      s1.add(f);
      s2.add(f * f);
      fmx = fmax(fmx, f);
      nz += f != 0.0 ? 1.0 : 0.0;
    } else {
      for (int d = 0; d < D; ++d) bins[d] = 0;
    }
    // stage (f^2, bins) of the tile; lane d then adds all 64 events into dimension d
    wv[wave * 64 + lane] = f * f;
    for (int d = 0; d < D; ++d) wb[(wave * 64 + lane) * D + d] = (uint8_t)bins[d];
    // the wave's LDS writes complete before any lane reads them (and reads before the
    // next tile overwrites them); the trip count is wave-uniform, so no block barrier
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane < D) {
      double* h = hist + wave * HB + lane * NB;
      for (int j = 0; j < 64; ++j) h[wb[(wave * 64 + j) * D + lane]] += wv[wave * 64 + j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  // block reduction of the scalars in a fixed tree order
  __shared__ double red[ACC_BLOCK][NSCAL];
  red[threadIdx.x][0] = s1.s;
  red[threadIdx.x][1] = s1.c;
  red[threadIdx.x][2] = s2.s;
  red[threadIdx.x][3] = s2.c;
  red[threadIdx.x][4] = fmx;
  red[threadIdx.x][5] = nz;
  __syncthreads();
  for (int off = ACC_BLOCK / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      KSum a{red[threadIdx.x][0], red[threadIdx.x][1]}, b{red[threadIdx.x + off][0], red[threadIdx.x + off][1]};
      a.merge(b);
      KSum c{red[threadIdx.x][2], red[threadIdx.x][3]}, d{red[threadIdx.x + off][2], red[threadIdx.x + off][3]};
      c.merge(d);
      red[threadIdx.x][0] = a.s;
      red[threadIdx.x][1] = a.c;
      red[threadIdx.x][2] = c.s;
      red[threadIdx.x][3] = c.c;
      red[threadIdx.x][4] = fmax(red[threadIdx.x][4], red[threadIdx.x + off][4]);
      red[threadIdx.x][5] += red[threadIdx.x + off][5];
    }
    __syncthreads();
  }
  double* out = partials + (size_t)blockIdx.x * (NSCAL + 2 * HB);
  if (threadIdx.x < NSCAL) out[threadIdx.x] = red[0][threadIdx.x];
  for (int i = threadIdx.x; i < HB; i += ACC_BLOCK) {
    KSum k;
    for (int w = 0; w < nwave; ++w) k.add(hist[w * HB + i]);
    out[NSCAL + i] = k.s;
    out[NSCAL + HB + i] = k.c;
  }
}

// result: [sum f, sum f^2, max f, n_nonzero, 0, 0] then ndim*nbins histogram
__global__ void k_reduce(int nparts, int hb, const double* __restrict__ partials, double* __restrict__ result) {
  const int stride = NSCAL + 2 * hb;
  for (int i = threadIdx.x; i < hb + 1; i += blockDim.x) {
    if (i == hb) {
      KSum a, b;
      double mx = 0.0, nz = 0.0;
      for (int p = 0; p < nparts; ++p) {
        const double* q = partials + (size_t)p * stride;
        a.merge(KSum{q[0], q[1]});
        b.merge(KSum{q[2], q[3]});
        mx = fmax(mx, q[4]);
        nz += q[5];
      }
      result[0] = a.s + a.c;
      result[1] = b.s + b.c;
      result[2] = mx;
      result[3] = nz;
      result[4] = 0.0;
      result[5] = 0.0;
    } else {
      KSum k;
      for (int p = 0; p < nparts; ++p) {
        const double* q = partials + (size_t)p * stride;
        k.merge(KSum{q[NSCAL + i], q[NSCAL + hb + i]});
      }
      result[NSCAL + i] = k.s + k.c;
    }
  }
}

// One block per dimension; thread 0 rebins (NB <= 64 entries: serial is cheapest).
__global__ void k_refine(double* __restrict__ grid, const double* __restrict__ hist, int nbins, double alpha) {
  if (threadIdx.x != 0) return;
  const int d = blockIdx.x;
  double* edges = grid + d * (nbins + 1);
  const double* h = hist + d * nbins;
  double sm[MAXBINS], reg[MAXBINS], ne[MAXBINS + 1];
  if (nbins < 2) return;
  sm[0] = (h[0] + h[1]) / 2.0;
  sm[nbins - 1] = (h[nbins - 2] + h[nbins - 1]) / 2.0;
  for (int i = 1; i < nbins - 1; ++i) sm[i] = (h[i - 1] + h[i] + h[i + 1]) / 3.0;
  double sum = 0.0;
  for (int i = 0; i < nbins; ++i) sum += sm[i];
  if (!(sum > 0.0) || !isfinite(sum)) return;
  double rsum = 0.0;
  for (int i = 0; i < nbins; ++i) {
    const double r = sm[i] / sum;
    reg[i] = (r > 0.0 && r < 1.0) ? pow((1.0 - r) / log(1.0 / r), alpha) : 0.0;
    rsum += reg[i];
  }
  if (!(rsum > 0.0)) return;
  const double ave = rsum / nbins;
  ne[0] = 0.0;
  ne[nbins] = 1.0;
  int k = -1;
  double acc = 0.0;
  for (int j = 1; j < nbins; ++j) {
    while (acc < ave && k < nbins - 1) {
      ++k;
      acc += reg[k];
    }
    acc -= ave;
    const double lo = edges[k], hi = edges[k + 1];
    ne[j] = hi - (hi - lo) * acc / reg[k];
  }
  for (int j = 0; j <= nbins; ++j) edges[j] = ne[j];
}

static int make_gen(const mfx_ps_params* p, GenConst& g) {
  if (!p) return fail(-3, "null parameters%s", "");
  const int n = p->n_out;
  if (n < 2 || n > RAMBO_MAXN) return fail(-1, "n_out must be in [2, 8]%s", "");
  if (p->mapping < 0 || p->mapping > 3)
    return fail(-2, "mapping must be 0 (RAMBO), 1 (diet), 2 (hadronic) or 3 (hadronic, multi-channel)%s", "");
  const int need = p->mapping == 0 ? 4 * n : p->mapping == 1 ? 3 * n - 4 : p->mapping == 2 ? 4 * n + 2 : 3 * n - 2;
  if (p->ndim != need)
    return fail(-2, "ndim must be 4*n_out (RAMBO), 3*n_out-4 (diet), 4*n_out+2 (hadronic) or 3*n_out-2 (mapping 3)%s", "");
  if (p->ndim > MAXDIM || p->nbins < 2 || p->nbins > MAXBINS) return fail(-2, "ndim/nbins out of range%s", "");
  memset(&g, 0, sizeof(g));
  g.rc.n = n;
  g.rc.massive = 0;
  g.rc.total_mass = 0.0;
  for (int i = 0; i < n; ++i) {
    g.rc.masses[i] = p->masses[i];
    g.rc.total_mass += p->masses[i];
    if (p->masses[i] != 0.0) g.rc.massive = 1;
  }
  if (g.rc.total_mass >= p->sqrts) return fail(-2, "not enough energy for the outgoing masses%s", "");
  if (p->mapping >= 2 && !(p->shat_min >= g.rc.total_mass * g.rc.total_mass && p->shat_min < p->sqrts * p->sqrts))
    return fail(-2, "hadronic mapping: shat_min must lie in [(sum of masses)^2, s)%s", "");
  g.rc.logw0 = log(MFX_PI / 2.0) * (n - 1) - 2.0 * lgamma((double)(n - 1)) - log((double)(n - 1));
  g.rc.inorm = 1.0 / pow(2.0 * MFX_PI, 3 * n - 4);
  g.ndim = p->ndim;
  g.nbins = p->nbins;
  g.mapping = p->mapping;
  g.nexternal = n + 2;
  g.sqrts = p->sqrts;
  g.norm = p->gev2pb / (2.0 * p->sqrts * p->sqrts);
  g.gev2pb = p->gev2pb;
  g.shat_min = p->shat_min;
  g.lab_frame = p->lab_frame;
  g.pt_min = p->pt_min;
  g.abs_eta_max = p->abs_eta_max;
  g.dr_min = p->dr_min;
  g.cut_first = p->cut_first_leg - 2;
  g.cut_last = n - 1;
  g.seed = p->seed;
  if (p->cut_first_leg < 2 || p->cut_first_leg > n + 2) return fail(-2, "cut_first_leg out of range%s", "");
  if (p->ncuts < 0 || p->ncuts > MFX_MAX_CUTS) return fail(-2, "ncuts must lie in [0, MFX_MAX_CUTS]%s", "");
  g.ncuts = p->ncuts;
  for (int c = 0; c < p->ncuts; ++c) {
    mfx_cut k = p->cuts[c];
    if (k.variable < 0 || k.variable > 2) return fail(-2, "cut variable must be 0 (pt), 1 (mt) or 2 (mt2)%s", "");
    if (k.particle < 2 || k.particle > n + 1) return fail(-2, "cuts apply to outgoing legs 2 .. n_out + 1%s", "");
    if (!k.use_min && !k.use_max) return fail(-2, "a cut needs a lower or an upper bound%s", "");
    k.particle -= 2;                       // outgoing-leg index of the momenta pass_cuts sees
    g.cuts[c] = k;
  }
  return 0;
}

}  // namespace mfx

using namespace mfx;

extern "C" {

int64_t mfx_integrand_workspace_size(int64_t nev, int ndim, int nbins) {
  const int64_t nblk = (nev + GEN_BLOCK - 1) / GEN_BLOCK;
  int64_t bytes = 0;
  bytes += (nblk + 1) * 4 * 2;                                  // counts, offsets
  bytes = (bytes + 255) / 256 * 256;
  bytes += (int64_t)ACC_GRID * (NSCAL + 2 * ndim * nbins) * 8;  // accumulate partials
  return bytes + 256;
}

static int ps_generate(const mfx_ps_params* prm, const double* grid, uint64_t ev0, int64_t nev, uint32_t iteration,
                       const float* y_in, const double* w_in, const double* x_in, double* all_ps, double* wps,
                       uint8_t* flag, int* slot, double* dense, double* x12, void* workspace, int64_t workspace_bytes,
                       int64_t* npass_host, void* stream) {
  GenConst g;
  int rc = make_gen(prm, g);
  if (rc) return rc;
  if (g.mapping == 3) return fail(-2, "mapping 3 is the multi-channel phase space's (mfx_mc_generate)%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev > 0x7fffffff) return fail(-2, "nev must fit in 31 bits%s", "");
  if (workspace_bytes < mfx_integrand_workspace_size(nev, prm->ndim, prm->nbins))
    return fail(-8, "workspace too small (see mfx_integrand_workspace_size)%s", "");
  if (nev == 0) {
    if (npass_host) *npass_host = 0;
    return 0;
  }
  if ((!grid && !x_in) || !all_ps || !wps || !flag || !slot || !dense || !workspace)
    return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((nev + GEN_BLOCK - 1) / GEN_BLOCK);
  int* counts = (int*)workspace;
  int* offsets = counts + nblk + 1;
  hipLaunchKernelGGL(k_generate, dim3(nblk), dim3(GEN_BLOCK), 0, s, g, grid, ev0, nev, iteration, y_in, w_in, x_in,
                     all_ps, wps, flag, x12, counts);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nblk, offsets);
  hipLaunchKernelGGL(k_compact, dim3(nblk), dim3(GEN_BLOCK), 0, s, g.nexternal, nev, all_ps, flag, offsets, dense,
                     slot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_status(e);
  if (npass_host) {
    int tot = 0;
    e = hipMemcpyAsync(&tot, offsets + nblk, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    *npass_host = tot;
  }
  return 0;
}

int mfx_ps_generate(const mfx_ps_params* prm, const double* grid, uint64_t ev0, int64_t nev, uint32_t iteration,
                    const float* y_in, const double* w_in, double* all_ps, double* wps, uint8_t* flag, int* slot,
                    double* dense, double* x12, void* workspace, int64_t workspace_bytes, int64_t* npass_host,
                    void* stream) {
  return ps_generate(prm, grid, ev0, nev, iteration, y_in, w_in, nullptr, all_ps, wps, flag, slot, dense, x12,
                     workspace, workspace_bytes, npass_host, stream);
}

int mfx_ps_generate_xrand(const mfx_ps_params* prm, const double* xrand, int64_t nev, double* all_ps, double* wps,
                          uint8_t* flag, int* slot, double* dense, double* x12, void* workspace,
                          int64_t workspace_bytes, int64_t* npass_host, void* stream) {
  if (nev > 0 && !xrand) return fail(-3, "null pointer (xrand)%s", "");
  return ps_generate(prm, nullptr, 0, nev, 0, nullptr, nullptr, xrand, all_ps, wps, flag, slot, dense, x12, workspace,
                     workspace_bytes, npass_host, stream);
}

int mfx_vegas_accumulate(const mfx_ps_params* prm, uint64_t ev0, int64_t nev, uint32_t iteration, const float* y_in,
                         const int* slot, const double* me, const double* wps, const double* f_in, double* f_out,
                         void* workspace, int64_t workspace_bytes, double* result, void* stream) {
  GenConst g;
  int rc = make_gen(prm, g);
  if (rc) return rc;
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (workspace_bytes < mfx_integrand_workspace_size(nev, prm->ndim, prm->nbins))
    return fail(-8, "workspace too small (see mfx_integrand_workspace_size)%s", "");
  if (!result || !workspace) return fail(-3, "null pointer%s", "");
  if (!f_in && nev > 0 && (!slot || !me || !wps)) return fail(-3, "need f_in or (slot, me, wps)%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nblk = (nev + GEN_BLOCK - 1) / GEN_BLOCK;
  int64_t off = (nblk + 1) * 4 * 2;
  off = (off + 255) / 256 * 256;
  double* partials = (double*)((char*)workspace + off);
  const int hb = g.ndim * g.nbins;
  const int nwave = ACC_BLOCK / 64;
  const size_t shmem = (size_t)nwave * hb * 8 + nwave * 64 * 8 + (size_t)nwave * 64 * g.ndim;
  if (shmem + sizeof(double) * ACC_BLOCK * NSCAL > 60 * 1024) return fail(-2, "ndim*nbins too large for LDS%s", "");
  hipLaunchKernelGGL(k_accumulate, dim3(ACC_GRID), dim3(ACC_BLOCK), shmem, s, g, ev0, nev, iteration, y_in, slot, me,
                     wps, f_in, f_out, partials);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, ACC_GRID, hb, partials, result);
  return hip_status(hipGetLastError());
}

static int make_mc(const mfx_ps_params* prm, const GenConst& g, const mfx_mc_channel* chs, int nch,
                   const double* alpha, double scale, McConst& mc) {
  if (g.mapping != 1 && g.mapping != 3)
    return fail(-2, "multi-channel phase space needs the diet parameters (mapping 1, or 3 for hadronic)%s", "");
  if (nch < 1 || nch > MC_MAXCH) return fail(-2, "between 1 and 8 channels%s", "");
  if (!chs || !alpha) return fail(-3, "null pointer%s", "");
  if (!(scale > 0.0) || !isfinite(scale)) return fail(-2, "scale must be > 0%s", "");
  memset(&mc, 0, sizeof(mc));
  mc.nch = nch;
  mc.scale = scale;
  const int n = g.rc.n;
  for (int c = 0; c < nch; ++c) {
    const mfx_mc_channel& in = chs[c];
    McChannel& ch = mc.ch[c];
    if (!(alpha[c] >= 0.0) || !isfinite(alpha[c])) return fail(-2, "channel weights must be >= 0%s", "");
    mc.alpha[c] = alpha[c];
    ch.type = in.type;
    ch.a = in.a;
    ch.b = in.b;
    ch.lo = in.lo;
    ch.hi = in.hi;
    int m = 0;
    if (in.type == 0) {
      for (int i = 0; i < n; ++i) ch.legs[m++] = i;
    } else if (in.type == 1 || in.type == 2) {
      if (in.a < 0 || in.a >= n || prm->masses[in.a] != 0.0) return fail(-2, "channel leg a must be a massless outgoing leg%s", "");
      if (in.type == 2 && (in.b < 0 || in.b >= n || in.b == in.a || prm->masses[in.b] != 0.0))
        return fail(-2, "s-channel leg b must be another massless outgoing leg%s", "");
      if (!(in.lo > 0.0 && in.hi > in.lo) || !isfinite(in.hi)) return fail(-2, "channel range must be 0 < lo < hi%s", "");
      for (int i = 0; i < n; ++i)
        if (i != in.a && !(in.type == 2 && i == in.b)) ch.legs[m++] = i;
      if (in.type == 2) ch.legs[m++] = -1;     // the pair
      if (m < 2) return fail(-2, "a channel's sub-system needs two or more legs%s", "");
    } else {
      return fail(-2, "channel type must be 0 (flat), 1 (t-channel) or 2 (s-channel)%s", "");
    }
    ch.nsub = m;
    RamboConst& rc = ch.sub;
    rc.n = m;
    rc.massive = in.type == 2 ? 1 : 0;
    rc.total_mass = 0.0;
    for (int j = 0; j < m; ++j) {
      rc.masses[j] = ch.legs[j] >= 0 ? prm->masses[ch.legs[j]] : 0.0;
      rc.total_mass += rc.masses[j];
      if (rc.masses[j] != 0.0) rc.massive = 1;
    }
    rc.logw0 = log(MFX_PI / 2.0) * (m - 1) - 2.0 * lgamma((double)(m - 1)) - log((double)(m - 1));
    rc.inorm = 1.0 / pow(2.0 * MFX_PI, 3 * m - 4);
  }
  return 0;
}

static int mc_generate(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                       double scale, int channel, const double* grids, uint64_t ev0, int64_t nev, uint32_t iteration,
                       const float* y_in, const double* w_in, double* all_ps, double* wps, uint8_t* flag, int* slot,
                       double* dense, double* ratios, double* x12, void* workspace, int64_t workspace_bytes,
                       int64_t* npass_host, void* stream) {
  GenConst g;
  int rc = make_gen(prm, g);
  if (rc) return rc;
  McConst mc;
  rc = make_mc(prm, g, channels, nch, alpha, scale, mc);
  if (rc) return rc;
  if (channel < 0 || channel >= nch) return fail(-2, "channel out of range%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev > 0x7fffffff) return fail(-2, "nev must fit in 31 bits%s", "");
  if (workspace_bytes < mfx_integrand_workspace_size(nev, prm->ndim, prm->nbins))
    return fail(-8, "workspace too small (see mfx_integrand_workspace_size)%s", "");
  if (nev == 0) {
    if (npass_host) *npass_host = 0;
    return 0;
  }
  if (!grids || !all_ps || !wps || !flag || !slot || !dense || !ratios || !workspace) return fail(-3, "null pointer%s", "");
  if (g.mapping == 3 && !x12) return fail(-3, "null pointer (x12, needed in hadronic mode)%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((nev + GEN_BLOCK - 1) / GEN_BLOCK);
  int* counts = (int*)workspace;
  int* offsets = counts + nblk + 1;
  hipLaunchKernelGGL(k_mc_generate, dim3(nblk), dim3(GEN_BLOCK), 0, s, g, mc, channel, grids, ev0, nev, iteration,
                     y_in, w_in, all_ps, wps, flag, ratios, (float*)nullptr, (double*)nullptr, x12, counts);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nblk, offsets);
  hipLaunchKernelGGL(k_compact, dim3(nblk), dim3(GEN_BLOCK), 0, s, g.nexternal, nev, all_ps, flag, offsets, dense,
                     slot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_status(e);
  if (npass_host) {
    int tot = 0;
    e = hipMemcpyAsync(&tot, offsets + nblk, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    *npass_host = tot;
  }
  return 0;
}

int mfx_mc_generate(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                    double scale, int channel, const double* grids, uint64_t ev0, int64_t nev, uint32_t iteration,
                    double* all_ps, double* wps, uint8_t* flag, int* slot, double* dense, double* ratios,
                    double* x12, void* workspace, int64_t workspace_bytes, int64_t* npass_host, void* stream) {
  return mc_generate(prm, channels, nch, alpha, scale, channel, grids, ev0, nev, iteration, nullptr, nullptr, all_ps,
                     wps, flag, slot, dense, ratios, x12, workspace, workspace_bytes, npass_host, stream);
}

int mfx_mc_generate_points(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                           double scale, int channel, const double* grids, uint64_t ev0, int64_t nev,
                           uint32_t iteration, const float* y_in, const double* w_in, double* all_ps, double* wps,
                           uint8_t* flag, int* slot, double* dense, double* ratios, double* x12, void* workspace,
                           int64_t workspace_bytes, int64_t* npass_host, void* stream) {
  if (nev > 0 && !y_in) return fail(-3, "null pointer (y_in)%s", "");
  return mc_generate(prm, channels, nch, alpha, scale, channel, grids, ev0, nev, iteration, y_in, w_in, all_ps, wps,
                     flag, slot, dense, ratios, x12, workspace, workspace_bytes, npass_host, stream);
}

int mfx_mc_generate_learned(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                            double scale, int channel, const double* grids, uint64_t ev0, int64_t nev,
                            uint32_t iteration, const float* y_in, double* all_ps, double* wps, uint8_t* flag,
                            int* slot, double* dense, double* ratios, float* y0_out, double* gdens, void* workspace,
                            int64_t workspace_bytes, int64_t* npass_host, void* stream) {
  GenConst g;
  int rc = make_gen(prm, g);
  if (rc) return rc;
  McConst mc;
  rc = make_mc(prm, g, channels, nch, alpha, scale, mc);
  if (rc) return rc;
  if (nch + 1 > MC_MAXCH) return fail(-2, "at most 7 channels beside the learned one%s", "");
  if (g.mapping == 3) return fail(-2, "the learned channel runs at fixed energy (mapping 1)%s", "");
  if (mc.ch[0].type != 0) return fail(-2, "the learned channel lives on channel 0's cube, which must be flat%s", "");
  if (channel < 0 || channel >= nch) return fail(-2, "channel out of range%s", "");
  if (y_in && channel != 0) return fail(-2, "learned points are mapped through channel 0%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev > 0x7fffffff) return fail(-2, "nev must fit in 31 bits%s", "");
  if (workspace_bytes < mfx_integrand_workspace_size(nev, prm->ndim, prm->nbins))
    return fail(-8, "workspace too small (see mfx_integrand_workspace_size)%s", "");
  if (nev == 0) {
    if (npass_host) *npass_host = 0;
    return 0;
  }
  if (!grids || !all_ps || !wps || !flag || !slot || !dense || !ratios || !y0_out || !gdens || !workspace)
    return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((nev + GEN_BLOCK - 1) / GEN_BLOCK);
  int* counts = (int*)workspace;
  int* offsets = counts + nblk + 1;
  hipLaunchKernelGGL(k_mc_generate, dim3(nblk), dim3(GEN_BLOCK), 0, s, g, mc, channel, grids, ev0, nev, iteration,
                     y_in, (const double*)nullptr, all_ps, wps, flag, ratios, y0_out, gdens, (double*)nullptr, counts);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nblk, offsets);
  hipLaunchKernelGGL(k_compact, dim3(nblk), dim3(GEN_BLOCK), 0, s, g.nexternal, nev, all_ps, flag, offsets, dense,
                     slot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_status(e);
  if (npass_host) {
    int tot = 0;
    e = hipMemcpyAsync(&tot, offsets + nblk, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    *npass_host = tot;
  }
  return 0;
}

int mfx_mc_combine(int64_t nev, int nch, double alpha_learned, const float* lq, const double* gdens, double* wps,
                   double* ratios, void* stream) {
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nch < 1 || nch + 1 > MC_MAXCH) return fail(-2, "between 1 and 7 channels beside the learned one%s", "");
  if (!(alpha_learned >= 0.0) || !isfinite(alpha_learned)) return fail(-2, "alpha_learned must be >= 0%s", "");
  if (nev == 0) return 0;
  if (!lq || !gdens || !wps || !ratios) return fail(-3, "null pointer%s", "");
  const unsigned nblk = (unsigned)((nev + GEN_BLOCK - 1) / GEN_BLOCK);
  hipLaunchKernelGGL(k_mc_combine, dim3(nblk), dim3(GEN_BLOCK), 0, (hipStream_t)stream, nev, nch, alpha_learned, lq,
                     gdens, wps, ratios);
  return hip_status(hipGetLastError());
}

int mfx_mc_moments(int64_t nev, int nch, const int* slot, const double* me, const double* wps, const double* ratios,
                   void* workspace, int64_t workspace_bytes, double* result, void* stream) {
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nch < 1 || nch > MC_MAXCH) return fail(-2, "between 1 and 8 channels%s", "");
  if (workspace_bytes < (int64_t)ACC_GRID * nch * 2 * 8) return fail(-8, "workspace too small%s", "");
  if (!result || !workspace || (nev > 0 && (!slot || !wps || !ratios))) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  double* partials = (double*)workspace;
  hipLaunchKernelGGL(k_mc_moments, dim3(ACC_GRID), dim3(ACC_BLOCK), 0, s, nev, nch, slot, me, wps, ratios, partials);
  hipLaunchKernelGGL(k_mc_reduce, dim3(1), dim3(64), 0, s, ACC_GRID, nch, partials, result);
  return hip_status(hipGetLastError());
}

int mfx_vegas_refine(double* grid, const double* hist, int ndim, int nbins, double alpha, void* stream) {
  if (ndim < 1 || ndim > MAXDIM || nbins < 2 || nbins > MAXBINS) return fail(-2, "ndim/nbins out of range%s", "");
  if (!grid || !hist) return fail(-3, "null pointer%s", "");
  hipLaunchKernelGGL(k_refine, dim3(ndim), dim3(64), 0, (hipStream_t)stream, grid, hist, nbins, alpha);
  return hip_status(hipGetLastError());
}

int64_t mfx_unweight_workspace_size(int64_t n) {
  const int64_t nblk = (n + GEN_BLOCK - 1) / GEN_BLOCK;
  return (2 * nblk + 1) * 4 + 256;
}

int mfx_unweight(const double* w, int64_t n, double wmax, uint32_t seed, uint32_t iteration, uint64_t ev0,
                 uint8_t* keep, int* index_out, void* workspace, int64_t workspace_bytes, int64_t* nkeep_host,
                 void* stream) {
  if (n < 0 || n > 0x7fffffff) return fail(-2, "n must be in [0, 2^31)%s", "");
  if (!(wmax > 0.0)) return fail(-2, "wmax must be > 0%s", "");
  if (workspace_bytes < mfx_unweight_workspace_size(n)) return fail(-8, "workspace too small (see mfx_unweight_workspace_size)%s", "");
  if (n == 0) {
    if (nkeep_host) *nkeep_host = 0;
    return 0;
  }
  if (!w || !keep || !index_out || !workspace) return fail(-3, "null pointer%s", "");
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((n + GEN_BLOCK - 1) / GEN_BLOCK);
  int* counts = (int*)workspace;
  int* offsets = counts + nblk;
  hipLaunchKernelGGL(k_unweight_flag, dim3(nblk), dim3(GEN_BLOCK), 0, s, w, n, wmax, ev0, seed, iteration, keep,
                     counts);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nblk, offsets);
  hipLaunchKernelGGL(k_unweight_compact, dim3(nblk), dim3(GEN_BLOCK), 0, s, keep, n, offsets, index_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_status(e);
  if (nkeep_host) {
    int tot = 0;
    e = hipMemcpyAsync(&tot, offsets + nblk, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    *nkeep_host = tot;
  }
  return 0;
}

}  // extern "C"
