// C ABI: RAMBO and Philox uniforms.  Declarations: include/madflow_amd.h
#include "mfx_phasespace.h"
#include "mfx_rng.h"
#include "mfx_api_common.h"
#include "../../include/madflow_amd.h"
#include <math.h>
#include <stdio.h>

namespace mfx {

RamboConst make_rambo_const(int n, const double* masses) {
  RamboConst rc;
  rc.n = n;
  rc.massive = 0;
  rc.total_mass = 0.0;
  for (int i = 0; i < RAMBO_MAXN; ++i) rc.masses[i] = 0.0;
  if (masses)
    for (int i = 0; i < n; ++i) {
      rc.masses[i] = masses[i];
      rc.total_mass += masses[i];
      if (masses[i] != 0.0) rc.massive = 1;
    }
  rc.logw0 = log(MFX_PI / 2.0) * (n - 1) - 2.0 * lgamma((double)(n - 1)) - log((double)(n - 1));
  rc.inorm = 1.0 / pow(2.0 * MFX_PI, 3 * n - 4);
  return rc;
}

// status per event: 0 physical, 1 not enough energy (sum of masses >= sqrts: zero momenta
// and weight, the per-event form of phasespace.py:152-159), 2 the massive Newton
// iteration did not converge (weight 0).
__global__ void k_rambo(const double* __restrict__ xrand, int64_t nev, int xstride, RamboConst rc,
                        const double* __restrict__ sqrts_ev, double sqrts, double* __restrict__ p_out,
                        double* __restrict__ wt_out, int* __restrict__ status) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  double p[RAMBO_MAXN][4];
  const double s = sqrts_ev ? sqrts_ev[e] : sqrts;
  double w = 0.0;
  int st = 0;
  if (!(s > rc.total_mass)) {
    st = 1;
    for (int i = 0; i < rc.n; ++i)
      for (int k = 0; k < 4; ++k) p[i][k] = 0.0;
  } else {
    bool converged;
    w = rambo_event(rc, xrand + e * xstride, s, p, converged);
    if (!converged) {
      st = 2;
      w = 0.0;
    }
  }
  for (int i = 0; i < rc.n; ++i)
    for (int k = 0; k < 4; ++k) p_out[(e * rc.n + i) * 4 + k] = p[i][k];
  wt_out[e] = w;
  if (status) status[e] = st;
}

__global__ void k_philox(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration,
                         double* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  for (int j = 0; 2 * j < ndim; ++j) {
    double u0, u1;
    philox_pair(ev0 + e, (uint32_t)j, seed, iteration, u0, u1);
    out[e * ndim + 2 * j] = u0;
    if (2 * j + 1 < ndim) out[e * ndim + 2 * j + 1] = u1;
  }
}

// the same uniforms rounded to fp32 (round to nearest), for consumers that work in fp32
__global__ void k_philox_f32(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration,
                             float* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  for (int j = 0; 2 * j < ndim; ++j) {
    double u0, u1;
    philox_pair(ev0 + e, (uint32_t)j, seed, iteration, u0, u1);
    out[e * ndim + 2 * j] = (float)u0;
    if (2 * j + 1 < ndim) out[e * ndim + 2 * j + 1] = (float)u1;
  }
}

}  // namespace mfx

using namespace mfx;

extern "C" {

int mfx_rambo(const double* xrand, int64_t nev, int xstride, int n, const double* sqrts_ev, double sqrts,
              const double* masses, double* p_out, double* wt_out, int* status, void* stream) {
  if (n < 2 || n > RAMBO_MAXN) return fail(-1, "RAMBO needs 2 <= n <= 8 outgoing particles%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (xstride < 4 * n) return fail(-2, "xstride must be >= 4 n%s", "");
  if (masses)
    for (int i = 0; i < n; ++i)
      if (!(masses[i] >= 0.0)) return fail(-2, "masses must be >= 0%s", "");
  RamboConst rc = make_rambo_const(n, masses);
  if (!sqrts_ev && !(sqrts > rc.total_mass)) {
    char msg[160];
    snprintf(msg, sizeof(msg), "not enough energy (sqrt(s) = %g) for outgoing masses summing to %g", sqrts,
             rc.total_mass);
    return fail(-9, "%s", msg);
  }
  if (nev == 0) return 0;
  if (!xrand || !p_out || !wt_out) return fail(-3, "null pointer%s", "");
  dim3 grid((unsigned)((nev + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_rambo, grid, blk, 0, (hipStream_t)stream, xrand, nev, xstride, rc, sqrts_ev, sqrts, p_out,
                     wt_out, status);
  return hip_status(hipGetLastError());
}

int mfx_philox_uniforms(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, double* out,
                        void* stream) {
  if (nev < 0 || ndim < 1) return fail(-2, "nev must be >= 0 and ndim >= 1%s", "");
  if (nev == 0) return 0;
  if (!out) return fail(-3, "null pointer%s", "");
  dim3 grid((unsigned)((nev + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_philox, grid, blk, 0, (hipStream_t)stream, ev0, nev, ndim, seed, iteration, out);
  return hip_status(hipGetLastError());
}

int mfx_philox_uniforms_f32(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, float* out,
                            void* stream) {
  if (nev < 0 || ndim < 1) return fail(-2, "nev must be >= 0 and ndim >= 1%s", "");
  if (nev == 0) return 0;
  if (!out) return fail(-3, "null pointer%s", "");
  dim3 grid((unsigned)((nev + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_philox_f32, grid, blk, 0, (hipStream_t)stream, ev0, nev, ndim, seed, iteration, out);
  return hip_status(hipGetLastError());
}

}  // extern "C"
