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
/* Compiled-in processes (g g > t t~ + 0..3 g).  A process name as in matrix_1_<name>.py
 * selects its default kernels (the faster form: colour-ordered for g g > t t~ g g and
 * g g > t t~ g g g, fused Feynman-diagram programs for the others); <name>_fused selects
 * the fused kernel of a process whose default is colour-ordered (gg_ttxgg_fused). */
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
 *   all_ps      (nev, nexternal, 4) f64 device, (E, px, py, pz), 16-byte aligned
 *   mass_width  host [mdl_MT, mdl_WT]
 *   gc          host [3][2] frozen couplings GC_10, GC_11, GC_12 (re, im); used when gc_ev is NULL
 *   gc_ev       NULL or device [3][nev][2] per-event couplings (running alpha_s)
 *   out         (nev,) f64 device
 *   precision   64 (FP64 build) or 32 (FP32 wavefunctions/amplitudes/colour sum)
 *   ref_compat  1: sqrt(1/2) rounded through float32 like the reference (wavefunctions_flow.py:10)
 *   workspace   device scratch owned by the caller, 256-byte aligned (NULL/0 for the fused
 *               processes).  Processes evaluated by the colour-ordered kernels stage their
 *               jamps there; any size of at least one 16-event tile works, the library
 *               splits events and helicity rows to fit.
 */
int mfx_smatrix(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                const double* gc_ev, double* out, int precision, int ref_compat, void* workspace,
                int64_t workspace_bytes, void* stream);

/* Bytes of workspace for one pass over nev events (per_helicity = 1: for mfx_matrix),
 * nev rounded up to a 16-event tile; 0 when the process needs none.  No counterpart in the reference: its TF op allocated
 * its intermediates itself (custom_op/generation.py:206-211). */
int64_t mfx_smatrix_workspace_size(int proc, int64_t nev, int precision, int per_helicity);

/* mfx_smatrix without a caller workspace, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference.  It allocates nothing: the colour-ordered processes
 * stage their jamps in a fixed device array of mfx_smatrix_noalloc_scratch_bytes()
 * (256 MiB) compiled into the library, one copy per GPU, held from the library's first
 * use on that GPU until the process ends; events and helicity rows are split into passes
 * that fit it, so any nev is accepted (a process whose single 16-event tile of one
 * helicity row would not fit is rejected with -8).  The fused processes use no scratch.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from several threads) race and give wrong results.  Arguments,
 * results and errors are otherwise those of mfx_smatrix. */
int mfx_smatrix_noalloc(int proc, const double* all_ps, int64_t nev, const double* mass_width, const double* gc,
                        const double* gc_ev, double* out, int precision, int ref_compat, void* stream);
int64_t mfx_smatrix_noalloc_scratch_bytes(void);

/* matrix(all_ps, hel, ...) of one helicity vector (matrix_method_python.inc:106-138):
 * colour-summed |M|^2 of that row (no averaging).  hel: host int[nexternal], entries
 * +-1, or 4 on a gluon for the BRST/gauge check (wavefunctions_flow.py:146-152).
 * jamp_out: NULL or device (nev, ncolor, 2) f64.  workspace as for mfx_smatrix. */
int mfx_matrix(int proc, const double* all_ps, int64_t nev, const int* hel, const double* mass_width,
               const double* gc, const double* gc_ev, double* out, double* jamp_out, int precision,
               int ref_compat, void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------ HELAS / ALOHA building blocks */
/* External wavefunction for nev momenta p (nev,4): kind 0 vxxxxx, 1 oxxxxx, 2 ixxxxx
 * (wavefunctions_flow.py:54-154), 3 sxxxxx (:32-51; nhel ignored, nsign = nss).
 * out: (nev, 6, 2) f64 device, or (nev, 3, 2) for the scalar. */
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
 * 4 uniforms per outgoing particle; masses host [n] (NULL = massless).  sqrts_ev NULL:
 * every event at the host value sqrts, rejected with -9 ("not enough energy") when the
 * masses sum to sqrts or more, like rambo(check_physical=True) (phasespace.py:152-159);
 * else per-event device values (nev,).  Massive events run their own Newton iteration
 * to |sum E - sqrts| <= 1e-14 sqrts (at most 50 steps).
 * p_out (nev, n, 4) f64 device, wt_out (nev,) f64 device.  status: NULL or device int
 * (nev,): 0 physical; 1 not enough energy for that event's sqrts (zero momenta and
 * weight); 2 the Newton iteration did not converge (weight 0). */
int mfx_rambo(const double* xrand, int64_t nev, int xstride, int n, const double* sqrts_ev, double sqrts,
              const double* masses, double* p_out, double* wt_out, int* status, void* stream);

/* Philox4x32-10 uniforms in (0,1): out (nev, ndim) f64, counter (ev0+e, dim/2, iteration), key (seed). */
int mfx_philox_uniforms(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, double* out,
                        void* stream);
/* The same uniforms rounded to fp32 (round to nearest). */
int mfx_philox_uniforms_f32(uint64_t ev0, int64_t nev, int ndim, uint32_t seed, uint32_t iteration, float* out,
                            void* stream);

/* ------------------------------------------------------ VEGAS integration */
/* One cut of PhaseSpaceGenerator.register_cut (phasespace.py:424-478): an event passes
 * when min_val < variable(particle) (if use_min) and variable(particle) < max_val (if
 * use_max), on the outgoing external leg `particle` (2 .. n_out + 1) in the partonic c.m.
 * frame, before any boost (phasespace.py:502-517).  variable: 0 pt (phasespace.py:416-421),
 * 1 mt = sqrt(mt2), 2 mt2 = m^2 + pt^2 (phasespace.py:405-414). */
#define MFX_MAX_CUTS 16
typedef struct mfx_cut {
  int variable, particle, use_min, use_max;
  double min_val, max_val;
} mfx_cut;

/* Phase-space / integrand configuration (plain C struct). */
typedef struct mfx_ps_params {
  int n_out;            /* outgoing particles (t, t~, gluons) */
  int ndim;             /* VEGAS dimensions: 4*n_out (RAMBO), 3*n_out-4 (RAMBO on diet), 4*n_out+2 (hadronic),
                           3*n_out-2 (mapping 3) */
  int nbins;            /* VEGAS bins per dimension (<= 64) */
  int mapping;          /* 0 RAMBO (phasespace.py:145), 1 RAMBO on diet (minimal dimensions),
                           2 hadronic: x1, x2 then RAMBO at sqrt(x1 x2 s) (ramboflow, phasespace.py:255-317),
                           3 hadronic multi-channel (mfx_mc_generate only): x1, x2 as in mapping 2, then the
                           channels' maps at sqrt(x1 x2 s) on the other 3*n_out-4 dimensions */
  double sqrts;         /* partonic sqrt(s) in GeV (mappings 0, 1) or the hadronic sqrt(s) (mappings 2, 3) */
  double masses[8];     /* outgoing masses */
  double gev2pb;        /* unit conversion applied with the flux 1/(2 s) (phasespace.py:315-317) */
  double pt_min;        /* cuts on outgoing legs >= cut_first_leg: pT > pt_min */
  double abs_eta_max;   /*   |eta| < abs_eta_max (0 disables) */
  double dr_min;        /*   Delta R(i,j) > dr_min between those legs (0 disables) */
  int cut_first_leg;    /* first external leg index the cuts apply to (n_out + 2: no cuts) */
  uint32_t seed;        /* Philox key */
  double shat_min;      /* mappings 2, 3: lower end of shat (phasespace.py:270-273 uses (sum of masses)^2) */
  int lab_frame;        /* mappings 2, 3: 1 = momenta boosted to the lab frame (phasespace.py:322-357; what the
                           reference's integration script writes, com_output = False at madflow_exec.py:389),
                           0 = the partonic c.m. frame (PhaseSpaceGenerator's own default, phasespace.py:377) */
  int reserved;
  int ncuts;            /* register_cut cuts in `cuts`, ANDed with the ones above (0 .. MFX_MAX_CUTS) */
  int reserved2;
  mfx_cut cuts[MFX_MAX_CUTS];
} mfx_ps_params;

/* Device workspace needed by mfx_ps_generate / mfx_vegas_accumulate for nev events. */
int64_t mfx_integrand_workspace_size(int64_t nev, int ndim, int nbins);

/* One VEGAS iteration's events: Philox uniforms (counter ev0+e, iteration), or the
 * points y_in (nev, ndim) f32 of an external proposal, -> VEGAS map (grid: device
 * [ndim][nbins+1]) -> RAMBO -> cuts (phasespace.py:424-520; in the partonic c.m. frame,
 * as the reference applies them).  Rejected with -2 before any launch when the outgoing
 * masses sum to sqrts or more (phasespace.py:152-159) or, for the hadronic mapping, when
 * shat_min lies outside [(sum of masses)^2, s).  w_in (optional, nev) multiplies the
 * weights (e.g. 1/q of the proposal).  Writes the full momenta all_ps (nev, n_out+2, 4),
 * wps = phase-space weight x VEGAS jacobian x gev2pb/(2 shat) (0 for cut events), flag
 * (1 = pass, 0 = cut, 2 = the massive Newton iteration did not converge), slot
 * (dense index or -1) and the passing events' momenta densely into dense (the input of
 * mfx_smatrix, the reference's boolean_mask compaction).  x12 (optional, nev x 2): the
 * momentum fractions x1, x2 (1, 1 for the fixed-energy mappings).  *npass_host receives
 * the number of passing events (this synchronises the stream). */
int mfx_ps_generate(const mfx_ps_params* prm, const double* grid, uint64_t ev0, int64_t nev, uint32_t iteration,
                    const float* y_in, const double* w_in, double* all_ps, double* wps, uint8_t* flag, int* slot,
                    double* dense, double* x12, void* workspace, int64_t workspace_bytes, int64_t* npass_host,
                    void* stream);
/* PhaseSpaceGenerator.__call__ / ramboflow (phasespace.py:255-317,480-520) on the caller's
 * uniforms: mfx_ps_generate with the points xrand (device, nev x ndim f64) used as they
 * are, with no VEGAS grid (jacobian 1); the same outputs, cuts and compaction.  With the
 * hadronic mapping (2) that is ramboflow: x1, x2 from the first two coordinates, RAMBO at
 * sqrt(x1 x2 s) on the others, the weight with the flux 1/(2 shat) and the GeV^-2 -> pb
 * factor, and the lab-frame boost after the cuts when lab_frame is set. */
int mfx_ps_generate_xrand(const mfx_ps_params* prm, const double* xrand, int64_t nev, double* all_ps, double* wps,
                          uint8_t* flag, int* slot, double* dense, double* x12, void* workspace,
                          int64_t workspace_bytes, int64_t* npass_host, void* stream);

/* Accumulate one iteration: f = me[slot] * wps (or f_in if given; the reference's
 * scatter_nd back to full length, madflow_exec.py:466-468).  result (device, 6 + ndim*nbins
 * doubles): [sum f, sum f^2, max f, #f != 0, 0, 0, histogram of f^2 per dimension and bin].
 * Sums are compensated FP64 in a fixed order (bitwise reproducible); the histogram bins
 * come from y_in if given, else from the Philox stream.  f_out: optional
 * per-event weights. */
int mfx_vegas_accumulate(const mfx_ps_params* prm, uint64_t ev0, int64_t nev, uint32_t iteration, const float* y_in,
                         const int* slot, const double* me, const double* wps, const double* f_in, double* f_out,
                         void* workspace, int64_t workspace_bytes, double* result, void* stream);

/* ------------------------------------------------ multi-channel phase space */
/* No counterpart in the reference (doc/source/usage.rst:84-99 names the limitation: RAMBO
 * and VEGAS only).  A channel maps the 3*n_out-4 dimensional cube onto the fixed-sqrt(s)
 * phase space (outgoing legs numbered from 0: t, t~, gluons):
 *   type 0 flat:       RAMBO on diet over every outgoing leg;
 *   type 1 t-channel:  massless leg a with pT ~ 1/pT^2 on [lo, hi] (hi = the largest
 *                      energy the leg can carry), rapidity and azimuth flat; the other
 *                      legs by diet in the rest frame of P - k_a;
 *   type 2 s-channel:  massless legs a, b with m_ab^2 ~ 1/m^2 on [lo, hi], isotropic in
 *                      their rest frame; the other legs and the pair by diet. */
typedef struct mfx_mc_channel {
  int type;
  int a, b;
  int reserved;
  double lo, hi;
} mfx_mc_channel;

/* mfx_ps_generate for the nev points of one channel (prm: the diet parameters, mapping 1):
 * Philox uniforms (counter ev0+e, iteration) -> VEGAS map of that channel's grid (grids:
 * device [nch][ndim][nbins+1]) -> the channel's map; wps = scale / sum_j alpha_j g_j(p)
 * x gev2pb / (2 s), with g_j the density of channel j at the point (its inverse map and
 * grid); ratios (device, nev x nch) receives g_j / g.  scale = alpha_channel * N / nev
 * makes the mean of f over all channels' points the stratified estimate.  The other
 * outputs, the workspace and *npass_host are those of mfx_ps_generate.
 * Mapping 3 (hadronic): the first two coordinates give x1, x2 as in mapping 2, the channels
 * act at sqrt(shat) with each channel's range capped by the kinematic limit there, wps
 * carries the (1 - tau_min)(-ln tau) weight and the flux 1 / (2 shat), the momenta are in
 * the lab frame when lab_frame is set (cuts in the partonic frame), and x12 (device,
 * nev x 2) receives x1, x2; x12 may be NULL at fixed energy. */
int mfx_mc_generate(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                    double scale, int channel, const double* grids, uint64_t ev0, int64_t nev, uint32_t iteration,
                    double* all_ps, double* wps, uint8_t* flag, int* slot, double* dense, double* ratios,
                    double* x12, void* workspace, int64_t workspace_bytes, int64_t* npass_host, void* stream);
/* mfx_mc_generate with the channel's cube points given (y_in, device (nev, ndim) f32; a
 * proposal on the cube such as a flow per channel) and their weights w_in (f64 (nev,),
 * 1 / the density of y on the cube, or NULL for 1): wps is multiplied by w_in, which keeps
 * the stratified estimate unbiased.  No reference counterpart. */
int mfx_mc_generate_points(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                           double scale, int channel, const double* grids, uint64_t ev0, int64_t nev,
                           uint32_t iteration, const float* y_in, const double* w_in, double* all_ps, double* wps,
                           uint8_t* flag, int* slot, double* dense, double* ratios, double* x12, void* workspace,
                           int64_t workspace_bytes, int64_t* npass_host, void* stream);

/* A learned channel beside the analytic ones (channel 0 must be flat): its density on the
 * phase space is q(y0(p)) g_0(p), with q the learned density on channel 0's cube and y0(p)
 * the point of that cube that channel 0's grid and map send to p.  With y_in (nev x ndim
 * f32, channel 0) the points are the learned channel's samples mapped through channel 0;
 * without, they are channel `channel`'s.  Writes y0_out (nev x ndim f32, the cube point
 * of every point, for the caller to evaluate q), the raw densities g_j in ratios (nev x
 * (nch + 1)), gdens (nev x 2: sum_j alpha_j g_j, g_0) and wps = scale x gev2pb / (2 s);
 * mfx_mc_combine then divides by the full density. */
int mfx_mc_generate_learned(const mfx_ps_params* prm, const mfx_mc_channel* channels, int nch, const double* alpha,
                            double scale, int channel, const double* grids, uint64_t ev0, int64_t nev,
                            uint32_t iteration, const float* y_in, double* all_ps, double* wps, uint8_t* flag,
                            int* slot, double* dense, double* ratios, float* y0_out, double* gdens, void* workspace,
                            int64_t workspace_bytes, int64_t* npass_host, void* stream);

/* g = gdens[e][0] + alpha_learned exp(lq[e]) gdens[e][1]: wps[e] /= g, ratios[e][j] /= g for
 * the nch analytic channels and ratios[e][nch] = exp(lq[e]) gdens[e][1] / g. */
int mfx_mc_combine(int64_t nev, int nch, double alpha_learned, const float* lq, const double* gdens, double* wps,
                   double* ratios, void* stream);

/* result (device, nch doubles) = sum over the nev points of f^2 ratios[e][j],
 * f = me[slot[e]] * wps[e] (0 when slot[e] < 0): the moments that adapt the channel
 * weights (Kleiss and Pittau).  Fixed-order compensated sums; workspace of at least
 * 1024 * nch * 16 bytes (mfx_integrand_workspace_size for the same nev suffices). */
int mfx_mc_moments(int64_t nev, int nch, const int* slot, const double* me, const double* wps, const double* ratios,
                   void* workspace, int64_t workspace_bytes, double* result, void* stream);

/* VEGAS grid refinement from the f^2 histogram: smoothing, ((1-r)/ln(1/r))^alpha damping
 * and equal-weight rebinning, per dimension (device grid updated in place). */
int mfx_vegas_refine(double* grid, const double* hist, int ndim, int nbins, double alpha, void* stream);

/* ------------------------------------------------ learned proposal (flow) */
/* Hit-or-miss unweighting of nev weighted events (replaces handing them to MadGraph's
 * lhe_parser unweight, lhe_writer.py:319-356): keep event e with probability
 * min(1, w[e] / wmax), the uniform taken from the Philox stream (ev0 + e, 0xffffffff,
 * iteration) with key (seed, 0x6d66).  keep (nev,) u8 device; index_out (nev,) i32 device
 * receives the kept event indices in order; *nkeep_host (optional, synchronises) their
 * number.  Events with w > wmax are kept with weight w / wmax, the others with weight 1. */
int64_t mfx_unweight_workspace_size(int64_t nev);
int mfx_unweight(const double* w, int64_t nev, double wmax, uint32_t seed, uint32_t iteration, uint64_t ev0,
                 uint8_t* keep, int* index_out, void* workspace, int64_t workspace_bytes, int64_t* nkeep_host,
                 void* stream);

/* Rational-quadratic spline coupling flow on [0,1]^D (no counterpart in the reference:
 * the north-star learned proposal).  masks: host int[L][D], 1 = dimension transformed
 * by that layer.  Parameters: one fp32 vector (sizes[0] floats) plus transposed bf16
 * copies of the hidden/output weights (sizes[1] elements) made by mfx_flow_cast.
 * sizes[] = {nparams, nbf16, NPmax, then per layer (offset, count)}. */
int mfx_flow_sizes(int D, int L, const int* masks, int64_t* sizes);
int mfx_flow_cast(int D, int L, const int* masks, const float* params, uint16_t* params_bf16, void* stream);
/* x -> u with the sum of the layers' log|du/dx| written to logdet (inverse = 0), or
 * u -> x with log|dx/du| (inverse = 1); one launch for all layers.  save: NULL, or
 * mfx_flow_save_bytes of device memory that receives what mfx_flow_backward needs
 * (forward direction only). */
int mfx_flow_apply(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16, int64_t n,
                   const float* x_in, float* x_out, float* logdet, int inverse, float* save, void* stream);
/* The forward pass of a training step with its minibatch drawn in the same launch: sample
 * e of the batch takes pool point k, the first with cdf[k] >= u_e (cdf: normalised running
 * sum of the npool training weights, f64), u_e the first 53-bit uniform of the Philox
 * stream (counter (e, 0, 0, draw), key (seed, 0x6d66)).  logdet (batch,) and save as in
 * mfx_flow_apply; y_out (batch, D) and f_out (batch,) optional (f_out needs pool_f).
 * guide (optional, nguide entries, mfx_flow_cdf_guide of the same cdf) finds the same k in
 * two or three loads instead of a bisection; NULL bisects.
 * No reference counterpart (the learned proposal is new). */
int mfx_flow_draw_forward(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16,
                          const double* cdf, const int* guide, int64_t nguide, int64_t npool, const float* pool_y,
                          const double* pool_f, uint32_t seed, uint32_t draw, int64_t batch, float* y_out,
                          double* f_out, float* logdet, float* save, void* stream);
int64_t mfx_flow_save_bytes(int D, int L, const int* masks, int64_t n);
int64_t mfx_flow_partial_bytes(int D, int L, const int* masks, int64_t n);
/* Gradient of sum_i gl_i * logdet_i (+ gx . u) w.r.t. the parameters (grad, fp32) and
 * the inputs (gx in/out; NULL: no gradient of the output, none returned), from a saved
 * forward pass.  partial: workspace of mfx_flow_partial_bytes; the reduction over samples
 * has a fixed order. */
int mfx_flow_backward(int D, int L, const int* masks, const float* params, const uint16_t* params_bf16, int64_t n,
                      const float* save, const float* gl, float* gx, float* partial, float* grad, void* stream);
/* Minibatch of a weighted training set: for each of the batch uniforms r, the first
 * index k with cdf[k] >= r (cdf: normalised running sum of the n weights, f64), whose
 * point y[k] (D floats) goes to yb and, if f is not NULL, f[k] to fb.  No reference
 * counterpart (the learned proposal is new). */
int mfx_flow_minibatch(const double* cdf, int64_t n, const double* r, int64_t batch, const float* y, int D,
                       const double* f, float* yb, double* fb, void* stream);
/* Backward seed gl_i = -v_i / sum_j v_j with v_i = w_i (f_i / q_i)^beta, q_i = exp(logq_i)
 * (w NULL: w_i = 1; beta = 0: KL, f unused), reduced in a fixed order; n <= 2^24. */
int mfx_flow_seed(const double* w, const double* f, const float* logq, int64_t n, float beta, float* gl, void* stream);
/* Adam update of the fp32 parameters (gradient clipped to +-clip, non-finite -> 0). */
int mfx_flow_adam(float* params, const float* grad, float* m1, float* m2, int64_t n, float lr, float beta1,
                  float beta2, float eps, int step, float clip, void* stream);
/* Normalised running sum of n training weights (f64; negative or non-finite weights count as
 * 0), the cdf that mfx_flow_minibatch and mfx_flow_draw_forward search; the sums run in a
 * fixed order.  workspace: mfx_flow_cdf_workspace(n) bytes of device memory, whose last
 * double receives the total (0: every weight is 0, and cdf is 0).  No reference counterpart. */
int64_t mfx_flow_cdf_workspace(int64_t n);
int mfx_flow_cdf(const double* w, int64_t n, double* cdf, double* workspace, int64_t workspace_bytes, void* stream);
/* Guide table of a cdf for mfx_flow_draw_forward: guide[b] (device int, m entries) = the
 * first index k with cdf[k] >= b / m.  With it a draw takes two or three dependent loads
 * instead of a bisection of log2(n) and returns the same index. */
int mfx_flow_cdf_guide(const double* cdf, int64_t n, int* guide, int64_t m, void* stream);
/* sum_i f_i^2 / ((1 - mix) exp(logq_i) + mix) in f64, fixed order: the second moment of the
 * importance weights of the defensive mixture at n points.  work: 257 doubles of device
 * memory; the sum lands in work[256]. */
int mfx_flow_moment(const float* logq, const double* f, int64_t n, double mix, double* work, void* stream);
/* winv_i = 1 / ((1 - mix) q_i + mix), f64 (n,), from the log-dets ld of mfx_flow_apply:
 * log q = ld after a forward pass (inverse = 0), -ld after an inverse pass (inverse = 1). */
int mfx_flow_mixture_winv(const float* ld, int64_t n, int inverse, double mix, double* winv, void* stream);
/* mfx_flow_adam on every layer's parameters followed by mfx_flow_cast's bf16 copies of the
 * updated weights, in one launch. */
int mfx_flow_adam_cast(int D, int L, const int* masks, float* params, uint16_t* params_bf16, const float* grad,
                       float* m1, float* m2, float lr, float beta1, float beta2, float eps, int step, float clip,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MADFLOW_AMD_H */
