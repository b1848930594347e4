// C ABI: process registry, smatrix/matrix entry points and the HELAS/ALOHA
// building-block kernels.  Declarations and reference citations: include/madflow_amd.h
#include "mfx_smatrix.h"
#include "mfx_api_common.h"
#include "../../include/madflow_amd.h"
#include <string.h>
#include <stdio.h>
#include <mutex>

namespace mfx {
static const int MAXPROC = 16;
struct Registry {
  ProcessEntry table[MAXPROC];
  int n = 0;
};
static Registry& registry() {
  static Registry r;  // function-local: safe whatever the order of static initialisers
  return r;
}
// Called once per process translation unit while the library loads; the table stays
// sorted by (number of external legs, name) so process indices do not depend on link order.
static bool registry_after(const ProcessEntry& a, const ProcessEntry& b) {
  return a.nexternal != b.nexternal ? a.nexternal > b.nexternal : strcmp(a.name, b.name) > 0;
}
int register_process(const ProcessEntry& e) {
  Registry& r = registry();
  if (r.n >= MAXPROC) return -1;
  int pos = r.n;
  while (pos > 0 && registry_after(r.table[pos - 1], e)) {
    r.table[pos] = r.table[pos - 1];
    --pos;
  }
  r.table[pos] = e;
  ++r.n;
  return pos;
}
static int nproc() { return registry().n; }
static const ProcessEntry& process(int i) { return registry().table[i]; }
#define NPROC nproc()

thread_local char g_last_error[512] = "";
int fail(int code, const char* fmt, const char* a) {
  snprintf(g_last_error, sizeof(g_last_error), fmt, a);
  return code;
}
int hip_status(hipError_t e) {
  if (e == hipSuccess) return 0;
  snprintf(g_last_error, sizeof(g_last_error), "HIP error: %s", hipGetErrorString(e));
  return (int)e;
}

// ---------------------------------------------------------- building-block kernels
template <class R>
__global__ void k_wavefunction(int kind, const double* __restrict__ p, int64_t nev, double mass, int nhel, int nsign,
                               double sqh, double* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  R q[4];
  for (int k = 0; k < 4; ++k) q[k] = R(p[4 * e + k]);
  W<R> w;
  if (kind == 0) vxxxxx(q, R(mass), R(nhel), R(nsign), R(sqh), w);
  else if (kind == 1) oxxxxx(q, R(mass), R(nhel), R(nsign), w);
  else if (kind == 2) ixxxxx(q, R(mass), R(nhel), R(nsign), w);
  else sxxxxx(q, R(nsign), w);
  const int n = kind == 3 ? 3 : 6;
  for (int k = 0; k < n; ++k) {
    out[(e * n + k) * 2] = double(w[k].re);
    out[(e * n + k) * 2 + 1] = double(w[k].im);
  }
}

template <class R>
__device__ __forceinline__ void load_w(const double* ins, int i, int64_t nev, int64_t e, W<R>& w) {
  for (int k = 0; k < 6; ++k) {
    const double* s = ins + ((i * nev + e) * 6 + k) * 2;
    w[k] = cx<R>(R(s[0]), R(s[1]));
  }
}

template <class R>
__global__ void k_aloha(int routine, const double* __restrict__ ins, int n_in, int64_t nev, double cre, double cim,
                        double M, double Wd, double* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  W<R> a, b, c, d, o;
  load_w(ins, 0, nev, e, a);
  if (n_in > 1) load_w(ins, 1, nev, e, b);
  if (n_in > 2) load_w(ins, 2, nev, e, c);
  if (n_in > 3) load_w(ins, 3, nev, e, d);
  const cx<R> cp{R(cre), R(cim)};
  const R m = R(M), wd = R(Wd);
  cx<R> amp(R(0), R(0));
  bool is_amp = false;
  switch (routine) {
    case 0: amp = FFV1_0(a, b, c, cp); is_amp = true; break;
    case 1: FFV1_1(a, b, cp, m, wd, o); break;
    case 2: FFV1_2(a, b, cp, m, wd, o); break;
    case 3: VVV1P0_1(a, b, cp, m, wd, o); break;
    case 4: VVVVP0_1s<1>(a, b, c, cp, m, wd, o); break;
    case 5: VVVVP0_1s<3>(a, b, c, cp, m, wd, o); break;
    case 6: VVVVP0_1s<4>(a, b, c, cp, m, wd, o); break;
    case 7: FFV1P0_3(a, b, cp, m, wd, o); break;
    case 8: amp = VVV1_0(a, b, c, cp); is_amp = true; break;
    case 9: amp = VVVV1_0(a, b, c, d, cp); is_amp = true; break;
    case 10: amp = VVVV3_0(a, b, c, d, cp); is_amp = true; break;
    default: amp = VVVV4_0(a, b, c, d, cp); is_amp = true; break;
  }
  if (is_amp) {
    out[2 * e] = double(amp.re);
    out[2 * e + 1] = double(amp.im);
  } else {
    for (int k = 0; k < 6; ++k) {
      out[(e * 6 + k) * 2] = double(o[k].re);
      out[(e * 6 + k) * 2 + 1] = double(o[k].im);
    }
  }
}

static double sqh_value(int ref_compat) {
  // reference: float_me(tf.math.sqrt(0.5)) — rounded through float32
  return ref_compat ? (double)(float)0.70710678118654752440 : 0.70710678118654752440;
}

static int fill_args(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                     const double* gc_ev, double* out, int precision, int ref_compat, SmatrixArgs& a) {
  if (proc < 0 || proc >= NPROC) return fail(-1, "unknown process index%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev > 0 && (!all_ps || !out)) return fail(-3, "null all_ps/out pointer%s", "");
  if (!mass_width || !gc) return fail(-4, "null mass_width/gc host pointer%s", "");
  if (precision != 32 && precision != 64) return fail(-5, "precision must be 32 or 64%s", "");
  if (((uintptr_t)all_ps) % 16 != 0) return fail(-6, "all_ps must be 16-byte aligned%s", "");
  memset(&a, 0, sizeof(a));
  a.all_ps = all_ps;
  a.nev = nev;
  a.mdl_MT = mass_width[0];
  a.mdl_WT = mass_width[1];
  for (int i = 0; i < 3; ++i) {
    a.gc[i][0] = gc[2 * i];
    a.gc[i][1] = gc[2 * i + 1];
  }
  a.gc_ev = gc_ev;
  a.sqh = sqh_value(ref_compat);
  a.out = out;
  return 0;
}

}  // namespace mfx

using namespace mfx;

extern "C" {

int mfx_version(void) { return 1; }
const char* mfx_last_error(void) { return g_last_error; }

int mfx_process_count(void) { return NPROC; }
const char* mfx_process_name(int proc) { return (proc >= 0 && proc < NPROC) ? process(proc).name : ""; }
int mfx_process_index(const char* name) {
  for (int i = 0; i < NPROC; ++i)
    if (name && strcmp(name, process(i).name) == 0) return i;
  return fail(-1, "unknown process '%s'", name ? name : "(null)");
}
int mfx_process_info(int proc, int* info) {
  if (proc < 0 || proc >= NPROC || !info) return fail(-1, "bad process index or info pointer%s", "");
  const ProcessEntry& p = process(proc);
  info[0] = p.nexternal;
  info[1] = p.ncomb;
  info[2] = p.ncolor;
  info[3] = p.ndiags;
  info[4] = p.ngraphs;
  return 0;
}

// workspace bytes for one pass over nev events (0: the process needs none)
int64_t mfx_smatrix_workspace_size(int proc, int64_t nev, int precision, int per_helicity) {
  if (proc < 0 || proc >= NPROC) return fail(-1, "unknown process index%s", "");
  if (precision != 32 && precision != 64) return fail(-5, "precision must be 32 or 64%s", "");
  const ProcessEntry& p = process(proc);
  if (!p.ws_per_row || nev <= 0) return 0;
  // passes are whole 16-event tiles (the jamp rows are tile-aligned)
  return (nev + 15) / 16 * 16 * (per_helicity ? 1 : p.ncomb) * p.ws_per_row(precision);
}

static int check_workspace(const ProcessEntry& p, int precision, int per_helicity, int64_t nev, const void* ws,
                           int64_t ws_bytes) {
  if (!p.ws_per_row || nev <= 0) return 0;
  const int64_t need = 16 * p.ws_per_row(precision);   // one 16-event tile of one helicity row
  if (!ws || ws_bytes < need) {
    char msg[192];
    snprintf(msg, sizeof(msg), "process %s needs a device workspace of at least %lld bytes "
             "(mfx_smatrix_workspace_size gives the size for one pass)", p.name, (long long)need);
    return fail(-8, "%s", msg);
  }
  if (((uintptr_t)ws) % 256 != 0) return fail(-6, "workspace must be 256-byte aligned%s", "");
  (void)per_helicity;
  return 0;
}

int mfx_smatrix(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                const double* gc_ev, double* out, int precision, int ref_compat, void* workspace,
                int64_t workspace_bytes, void* stream) {
  SmatrixArgs a;
  int rc = fill_args(proc, all_ps, nev, mass_width, gc, gc_ev, out, precision, ref_compat, a);
  if (rc) return rc;
  const ProcessEntry& p = process(proc);
  rc = check_workspace(p, precision, 0, nev, workspace, workspace_bytes);
  if (rc) return rc;
  return hip_status(p.launch(a, precision, 0, workspace, workspace_bytes, (hipStream_t)stream));
}

}  // extern "C"

namespace mfx {
// Device scratch of mfx_smatrix_noalloc: the jamp staging of the colour-ordered kernels,
// one copy per GPU, held from the library's first use on that GPU until the process ends.
// In host code the name is only a stand-in; its device address comes from
// hipGetSymbolAddress, once per device.
constexpr int64_t NOALLOC_SCRATCH_BYTES = 256ll << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static void* g_scratch_ptr[MAX_DEVICES] = {};
static std::mutex g_scratch_mutex;

// Device address of the current device's scratch array (looked up on the first call on
// that device); nullptr with mfx_last_error() set on failure.
static void* scratch_ptr(int* rc) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *rc = hip_status(e);
    return nullptr;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    *rc = fail(-10, "device index out of range for the scratch table%s", "");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *rc = e != hipSuccess ? hip_status(e) : fail(-10, "hipGetSymbolAddress returned no address%s", "");
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  *rc = 0;
  return g_scratch_ptr[dev];
}
}  // namespace mfx

extern "C" {

int64_t mfx_smatrix_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int mfx_smatrix_noalloc(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                        const double* gc_ev, double* out, int precision, int ref_compat, void* stream) {
  SmatrixArgs a;
  int rc = fill_args(proc, all_ps, nev, mass_width, gc, gc_ev, out, precision, ref_compat, a);
  if (rc) return rc;
  const ProcessEntry& p = process(proc);
  if (!p.ws_per_row || nev == 0)   // the fused kernels stage nothing
    return hip_status(p.launch(a, precision, 0, nullptr, 0, (hipStream_t)stream));
  // the launcher splits events and helicity rows to fit the scratch; it needs room for
  // one 16-event tile of one helicity row
  if (16 * p.ws_per_row(precision) > NOALLOC_SCRATCH_BYTES)
    return fail(-8, "process %s needs more jamp staging than the fixed scratch of mfx_smatrix_noalloc holds", p.name);
  void* ws = scratch_ptr(&rc);
  if (!ws) return rc;
  return hip_status(p.launch(a, precision, 0, ws, NOALLOC_SCRATCH_BYTES, (hipStream_t)stream));
}

int mfx_matrix(int proc, const double* all_ps, int64_t nev, const int* hel, const double* mass_width,
               const double* gc, const double* gc_ev, double* out, double* jamp_out, int precision,
               int ref_compat, void* workspace, int64_t workspace_bytes, void* stream) {
  SmatrixArgs a;
  int rc = fill_args(proc, all_ps, nev, mass_width, gc, gc_ev, out, precision, ref_compat, a);
  if (rc) return rc;
  if (!hel) return fail(-4, "null helicity vector%s", "");
  const ProcessEntry& p = process(proc);
  for (int l = 0; l < p.nexternal; ++l) {
    if (hel[l] != 1 && hel[l] != -1 && hel[l] != 4) return fail(-7, "helicities must be +-1 (or 4 on a gluon)%s", "");
    a.helv[l] = hel[l];
  }
  rc = check_workspace(p, precision, 1, nev, workspace, workspace_bytes);
  if (rc) return rc;
  a.use_helv = 1;
  a.jamp_out = jamp_out;
  return hip_status(p.launch(a, precision, 1, workspace, workspace_bytes, (hipStream_t)stream));
}

int mfx_wavefunction(int kind, const double* p, int64_t nev, double mass, int nhel, int nsign, double* out,
                     int precision, int ref_compat, void* stream) {
  if (kind < 0 || kind > 3) return fail(-1, "kind must be 0 (v), 1 (o), 2 (i) or 3 (s)%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev == 0) return 0;
  if (!p || !out) return fail(-3, "null pointer%s", "");
  dim3 grid((unsigned)((nev + 255) / 256)), blk(256);
  double sqh = sqh_value(ref_compat);
  if (precision == 32)
    hipLaunchKernelGGL(k_wavefunction<float>, grid, blk, 0, (hipStream_t)stream, kind, p, nev, mass, nhel, nsign, sqh,
                       out);
  else
    hipLaunchKernelGGL(k_wavefunction<double>, grid, blk, 0, (hipStream_t)stream, kind, p, nev, mass, nhel, nsign,
                       sqh, out);
  return hip_status(hipGetLastError());
}

int mfx_aloha(int routine, const double* ins, int n_in, int64_t nev, const double* coup, double mass, double width,
              double* out, int precision, void* stream) {
  static const int nin_needed[12] = {3, 2, 2, 2, 3, 3, 3, 2, 3, 4, 4, 4};
  if (routine < 0 || routine > 11) return fail(-1, "unknown ALOHA routine id%s", "");
  if (n_in != nin_needed[routine]) return fail(-2, "wrong number of input wavefunctions%s", "");
  if (nev < 0) return fail(-2, "nev must be >= 0%s", "");
  if (nev == 0) return 0;
  if (!ins || !out || !coup) return fail(-3, "null pointer%s", "");
  dim3 grid((unsigned)((nev + 255) / 256)), blk(256);
  if (precision == 32)
    hipLaunchKernelGGL(k_aloha<float>, grid, blk, 0, (hipStream_t)stream, routine, ins, n_in, nev, coup[0], coup[1],
                       mass, width, out);
  else
    hipLaunchKernelGGL(k_aloha<double>, grid, blk, 0, (hipStream_t)stream, routine, ins, n_in, nev, coup[0], coup[1],
                       mass, width, out);
  return hip_status(hipGetLastError());
}

}  // extern "C"
