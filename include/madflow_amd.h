/* madflow_amd — C ABI of the MI355X (gfx950) event-parallel hot path.
 *
 * Every buffer is caller-owned device memory (e.g. a torch tensor's data_ptr());
 * nothing here allocates in the hot path.  Calls are asynchronous on the given HIP
 * stream (pass torch.cuda.current_stream().cuda_stream, or NULL for the default
 * stream).  Return value: 0 on success, a negative code on a rejected argument
 * (the message is in mfx_last_error()), or a positive hipError_t from a launch.
 *
 * Reference interfaces replaced (paths relative to the madflow repository):
 *   - the generated process module's smatrix/matrix
 *       madgraph_plugin/template_files/matrix_method_python.inc:80-138
 *   - the TensorFlow custom op REGISTER_OP("Matrix<proc>") and its launch
 *       python_package/madflow/custom_op/constants.py:104-113, 164-198
 *   - HELAS external wavefunctions  python_package/madflow/wavefunctions_flow.py:32-154
 *   - ALOHA routines  python_package/madflow/tests/mockup_debug_me.py:34-397
 *   - RAMBO / cuts    python_package/madflow/phasespace.py:145-212, 424-520
 *   - the VEGAS integrand contract  python_package/madflow/scripts/madflow_exec.py:419-472
 */
#ifndef MADFLOW_AMD_H
#define MADFLOW_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- library */
int mfx_version(void);
/* Message of the last rejected call on this thread ("" if none). */
const char* mfx_last_error(void);

/* ------------------------------------------------------ process registry */
/* Compiled-in processes (g g > t t~ + 0..3 g); names as in matrix_1_<name>.py. */
int mfx_process_count(void);
const char* mfx_process_name(int proc);
int mfx_process_index(const char* name);
/* info[0..4] = nexternal, ncomb, ncolor, ndiags, ngraphs */
int mfx_process_info(int proc, int* info);

/* ------------------------------------------------------ matrix elements */
/* smatrix(all_ps, mdl_MT, mdl_WT, GC_10, GC_11[, GC_12]) -> f64[nev]
 * Replaces Matrix_<proc>.smatrix (matrix_method_python.inc:80-104) and the custom op
 * Matrix<proc> (custom_op/constants.py:104-113): sum over every helicity row of the
 * colour-summed |M|^2, divided by the process denominator.
 *   all_ps      (nev, nexternal, 4) f64 device, (E, px, py, pz)
 *   mass_width  host [mdl_MT, mdl_WT]
 *   gc          host [3][2] frozen couplings GC_10, GC_11, GC_12 (re, im); used when gc_ev is NULL
 *   gc_ev       NULL or device [3][nev][2] per-event couplings (running alpha_s)
 *   out         (nev,) f64 device
 *   precision   64 (FP64 build) or 32 (FP32 wavefunctions/amplitudes/colour sum)
 *   ref_compat  1: sqrt(1/2) rounded through float32 like the reference (wavefunctions_flow.py:10)
 */
int mfx_smatrix(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                const double* gc_ev, double* out, int precision, int ref_compat, void* stream);

/* matrix(all_ps, hel, ...) of one helicity vector (matrix_method_python.inc:106-138):
 * colour-summed |M|^2 of that row (no averaging).  hel: host int[nexternal], entries
 * +-1, or 4 on a gluon for the BRST/gauge check (wavefunctions_flow.py:146-152).
 * jamp_out: NULL or device (nev, ncolor, 2) f64. */
int mfx_matrix(int proc, const double* all_ps, int64_t nev, const int* hel, const double* mass_width,
               const double* gc, const double* gc_ev, double* out, double* jamp_out, int precision,
               int ref_compat, void* stream);

/* ------------------------------------------ HELAS / ALOHA building blocks */
/* External wavefunction for nev momenta p (nev,4): kind 0 vxxxxx, 1 oxxxxx, 2 ixxxxx
 * (wavefunctions_flow.py:54-154).  out: (nev, 6, 2) f64 device. */
int mfx_wavefunction(int kind, const double* p, int64_t nev, double mass, int nhel, int nsign, double* out,
                     int precision, int ref_compat, void* stream);

/* One ALOHA routine over nev input sets.  routine ids: 0 FFV1_0, 1 FFV1_1, 2 FFV1_2,
 * 3 VVV1P0_1, 4 VVVV1P0_1, 5 VVVV3P0_1, 6 VVVV4P0_1, 7 FFV1P0_3, 8 VVV1_0, 9 VVVV1_0,
 * 10 VVVV3_0, 11 VVVV4_0.  ins: device [n_in][nev][6][2]; coup host (re, im);
 * out: device (nev, 2) for amplitudes or (nev, 6, 2) for off-shell currents. */
int mfx_aloha(int routine, const double* ins, int n_in, int64_t nev, const double* coup, double mass,
              double width, double* out, int precision, void* stream);

/* --------------------------------------------------------- phase space */
/* RAMBO (phasespace.py:145-212) for nev events: xrand (nev, xstride) f64 device with
 * 4 uniforms per outgoing particle; sqrts per event (sqrts_stride 1) or one value
 * (sqrts_stride 0); masses host [n] (NULL = massless).  Massive events run their own
 * Newton iteration to |sum E - sqrts| <= 1e-14 sqrts.
 * p_out (nev, n, 4) f64 device, wt_out (nev,) f64 device. */
int mfx_rambo(const double* xrand, int64_t nev, int xstride, int n, const double* sqrts, int sqrts_stride,
              const double* masses, double* p_out, double* wt_out, void* stream);

/* Philox4x32-10 uniforms in (0,1): out (nev, ndim) f64, counter (ev0+e, dim/2, iteration), key (seed). */
int mfx_philox_uniforms(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, double* out,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MADFLOW_AMD_H */
