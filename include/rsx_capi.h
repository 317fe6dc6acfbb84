/*
 * rsx_capi.h -- C ABI of the MI355X Debye-SAXS engine (libresaxs_amd.so).
 *
 * Plain pointers and sizes only.  Every function returns an int status:
 * 0 on success, or one of the negative codes below (the first six are the
 * reference's resaxs::error codes, resaxslib/include/resaxs.hpp:213-221,
 * with the same values).  rsx_last_error() gives the message of the last
 * failure on an engine.
 *
 * Bodies are 16-byte {x, y, z, w} records (FP32) with the atom-type index
 * stored in the bit pattern of w, as the reference's cl_float4 bodies
 * (resaxslib/include/utils.hpp:61-88).  Factor tables are row-major
 * [n_q][n_types] (resaxslib/include/ff_coef.hpp:129-132).
 *
 * Threading: one engine owns one HIP stream; calls on an engine are
 * synchronous unless named *_async, and an engine must not be used from two
 * host threads at once (the reference's queue-per-algorithm rule).
 */
#ifndef RSX_CAPI_H
#define RSX_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RSX_OK                    0
#define RSX_INVALID_DEVICE       -1   /* resaxs::error::SAXS_INVALID_DEVICE */
#define RSX_ALG_NOT_INITIALIZED  -2   /* SAXS_ALG_NOT_INITIALIZED */
#define RSX_INVALID_ARG          -3   /* SAXS_INVALID_ARG */
#define RSX_PROGRAM_NOT_LOADED   -4   /* SAXS_PROGRAM_NOT_LOADED */
#define RSX_ARGS_NOT_SET         -5   /* SAXS_ARGS_NOT_SET */
#define RSX_NO_INITIAL_CALC      -6   /* SAXS_NO_INITIAL_CALC */
#define RSX_HIP_ERROR          -100   /* a HIP runtime call failed */
#define RSX_OUT_OF_MEMORY      -101
#define RSX_NOT_SUPPORTED      -102
#define RSX_PDB_FILE_NOT_FOUND -1100  /* parse_pdb.hpp SAXS_PDB_FILE_NOT_FOUND */

/* dot-sphere kinds for SASA */
#define RSX_DOTS_DODECA_ICOSA     0   /* 32 dots, rsasa_algorithm.cpp:157-208 (reference GPU default) */
#define RSX_DOTS_FIBONACCI        1   /* 2n dots, rsasa_algorithm.cpp:96-116 */

typedef struct rsx_engine rsx_engine;

/* ---- engine lifetime (replaces cl_base<FLT>::initialize, re_cl_algorithm_base.cpp:38-46) ---- */
int rsx_device_count(int *count);
int rsx_create(int device, rsx_engine **out);
int rsx_destroy(rsx_engine *h);
const char *rsx_last_error(const rsx_engine *h);
/* HIP stream of the engine (hipStream_t), for callers that order their own work with it */
void *rsx_stream(rsx_engine *h);

/* ---- parameters (saxs_params / form_factor_params / implicit_water_params,
 *      resaxslib/include/saxs_algorithm.hpp:33-132) ---- */
int rsx_set_q(rsx_engine *h, const float *q, int n_q);
/* table [n_q][n_types]; water_index < 0: no hydration factor */
int rsx_set_factors(rsx_engine *h, const float *table, int n_types, int water_index);
/* the default atomic factor table (atomic_form_factors::generate, ff_coef.hpp:112-211, as the
 * reference's engine sets it up, saxs_alg_cl_gpu_pt_wf.cpp:38-48): [n_q][16] at the engine's
 * q, mean atomic radius avg_radius (atomic_form_factors::calc_avg_radius of the bodies) and
 * the current expansion factor, water row OH2 (type 14) */
int rsx_set_atomic_factors(rsx_engine *h, float avg_radius);
/* excluded-volume expansion factor c > 0 (form_factor_params::set_expansion_factor,
 * saxs_algorithm.hpp:101-107; saxs_alg_cl_gpu_pt_wf.cpp:53-62).  With the atomic table of
 * rsx_set_atomic_factors the table is regenerated at c; a caller's table (rsx_set_factors) is
 * kept and c only recorded, as the reference's generic form-factor parameters do.  Either way
 * the next recalc is a full recompute (the parameters are dirty). */
int rsx_set_expansion_factor(rsx_engine *h, float c);
int rsx_set_water_weight(rsx_engine *h, float w);
/* skip the SASA recomputation on recalc once one exists (implicit_water_params::no_rsasa_recalc_) */
int rsx_set_no_rsasa_recalc(rsx_engine *h, int flag);
/* SASA set-up used for hydration: per-type radii, probe radius, dot sphere.
 * kind RSX_DOTS_FIBONACCI uses n_dots = 2*n_half; RSX_DOTS_DODECA_ICOSA ignores n_half. */
int rsx_set_sasa_params(rsx_engine *h, const float *radii, int n_types, float probe, int kind, int n_half);
/* full-recompute period of the running incremental I(q) (default 10000 updates; 0 = never).
 * No counterpart in the reference, which re-reduces its stored block sums on every update. */
int rsx_set_rebaseline(rsx_engine *h, int every_n_recalcs);
/* size the engine for n_bodies (saxs_cl_gpu::initialize + set_args, saxs_alg_cl_gpu_pt_wf.cpp:176-263) */
int rsx_initialize(rsx_engine *h, int n_bodies);

/* ---- Debye I(q) (i_saxs, saxs_algorithm.hpp:147-183) ---- */
/* saxs_cl_gpu::calc_curve, saxs_alg_cl_gpu_pt_wf.cpp:414-458 */
int rsx_calc_curve(rsx_engine *h, const float *bodies, int n, float *out_iq);
int rsx_calc_curve_f64(rsx_engine *h, const float *bodies, int n, double *out_iq);
/* saxs_cl_gpu::recalc_curve, saxs_alg_cl_gpu_pt_wf.cpp:468-577: bodies[start, start+len) moved */
int rsx_recalc_curve(rsx_engine *h, const float *bodies, int n, int start, int len, float *out_iq);
int rsx_recalc_curve_f64(rsx_engine *h, const float *bodies, int n, int start, int len, double *out_iq);
/* saxs_cl_gpu::commit / revert, saxs_alg_cl_gpu_pt_wf.cpp:592-620 */
int rsx_commit(rsx_engine *h);
int rsx_revert(rsx_engine *h);
/* exposed fraction per body from the last hydration update (implicit_water_params::access_rsasa) */
int rsx_get_rsasa(rsx_engine *h, float *out, int n);
/* engine state: 0 blank, 1 initialized, 2 program loaded, 3 args set, 4 computing (re_algorithm.hpp:34-54) */
int rsx_state(const rsx_engine *h);

/* ---- device-resident variants (inputs already in HBM; no host copies) ---- */
/* bodies_dev: device pointer to n float4 bodies in the caller's order.  Reuses the layout of
 * the last host calc_curve (slot order, jobs, path): RSX_NO_INITIAL_CALC before one,
 * RSX_ARGS_NOT_SET when parameters changed since.  The bodies are checked on the device:
 * a type outside the factor table is skipped, and (recurrence path) positions must stay
 * within min(1.25 R, 1.5 / dq) of the layout's centre, R its bounding radius, for the
 * recurrence's accuracy; a failed check is reported by the next rsx_synchronize. */
int rsx_calc_curve_dev_async(rsx_engine *h, const void *bodies_dev, int n, double *out_dev);
/* waits for the engine's stream; RSX_INVALID_ARG when a device-resident forward enqueued since
 * the last synchronize failed its checks (the message says which) */
int rsx_synchronize(rsx_engine *h);

/* ---- SASA (i_rsasa::calc_rsasa, rsasa_algorithm.hpp:42-61; kernel rsasa_dot_spheres.cl:94-163) ---- */
/* radii per type, dots [n_types][n_dots][4] templates relative to the atom centre;
 * out_counts exposed dots per body, out_frac = count / n_dots (either may be NULL) */
int rsx_calc_rsasa(rsx_engine *h, const float *bodies, int n, const float *radii, int n_types,
                   float probe, const float *dots, int n_dots, unsigned *out_counts, float *out_frac);

/* ---- SASA counts that allocate nothing (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc.  The exposed-dot counts of rsx_calc_rsasa on
 * device-resident data: bodies_dev n float4 {x, y, z, type bits}, dots_dev
 * [n_types][n_dots] float4 templates, counts_dev n unsigned (output); radii [n_types] on the
 * host (copied by value into the launch).  The cell list lives in a static device buffer of
 * RSX_SASA_NOALLOC_SCRATCH_BYTES on each GPU it runs on (held until the process ends), so
 * n <= RSX_SASA_NOALLOC_MAX_ATOMS, n_types <= 64 and 1 <= n_dots <= 512, else
 * RSX_INVALID_ARG with nothing launched.  A body whose type is not below n_types has no dots
 * (count 0) and occludes nothing.  Runs on the current HIP device, asynchronously on stream
 * (a hipStream_t of that device; NULL = the null stream); no hipMalloc, hipFree or memset.
 * Every call on one GPU shares its buffer: calls that overlap on one GPU (two streams or two
 * host threads) race. */
#define RSX_SASA_NOALLOC_MAX_ATOMS 65536
#define RSX_SASA_NOALLOC_SCRATCH_BYTES 3147264
int rsx_calc_rsasa_noalloc(const void *bodies_dev, int n, const float *radii, int n_types, float probe,
                           const void *dots_dev, int n_dots, unsigned *counts_dev, void *stream);

/* ---- host helpers (no device work) ---- */
/* rsasa_algorithm.cpp:96-116: 2*n_half dots {x,y,z,radius} on a sphere of radius+probe */
int rsx_fibonacci_dots(int n_half, float radius, float probe, float *out);
/* rsasa_algorithm.cpp:157-208: 32 dots {x,y,z,radius} */
int rsx_dodeca_icosa_dots(float radius, float probe, float *out);
/* rsasa_algorithm.cpp:58-80 (create_sphere_dots): rings of dots for n_sect sectors; writes at most
 * cap dots {x,y,z,radius} to out and the full count to *n_dots.  The reference's create_dots uses
 * n_sect = 2 pi radius sqrt(density), create_32_dots n_sect = 10.5. */
int rsx_sphere_dots(float radius, float n_sect, float probe, float *out, int cap, int *n_dots);
/* name, architecture, CU count and memory of a HIP device (report_opencl_caps, utils.hpp:38) */
int rsx_device_name(int device, char *buf, int len);
/* ff_coef.hpp:112-211: default table [n_q][16] and the 16 type radii */
int rsx_atomic_form_factors(const float *q, int n_q, float avg_radius, float expansion, float *out);
int rsx_atomic_radii(float *out16);

/* ---- distance-histogram Debye (approximate; no reference counterpart: the reference's
 * saxs_hist_cl_gpu class, saxs_algorithm.hpp:188-193, runs the exact pt_wf kernels) ---- */
/* I(q) [n_q] (FP64) of n bodies from their pair-distance histogram at bin width
 * bin_width (A), per type pair, with the engine's q, factor table and hydration
 * settings (SASA counts from the engine's dot spheres when the water weight is set).
 * Exact at q = 0; elsewhere each pair's sinc is taken at its bin centre: the error
 * falls as bin_width^2 (stated by the tests).  Needs rsx_set_q and rsx_set_factors, not
 * rsx_initialize; the engine's incremental state is untouched. */
int rsx_calc_curve_hist(rsx_engine *h, const float *bodies, int n, float bin_width, double *out);

/* ---- ensemble fit (calc_saxs/calc_saxs.hpp:420-624; the C++ driver is include/resaxs/calc_saxs.hpp) ---- */
/* Grid search of the expansion factor and water weight for the uniform mean of the
 * models' hydrated I(q) (default atomic factors, GPU SASA), least-squares scale and chi
 * against (q, i_exp, sigma).  Models: bodies [offsets[m], offsets[m+1]), all of one atom
 * count.  Each range is {lo, hi}; lo == hi fixes the value (for the scale: no fit).
 * out = {expansion factor, water weight, scale, chi, model evaluations}; out_iq (may be
 * NULL) = the scaled best intensity [n_q]. */
int rsx_fit_ensemble(const float *bodies, const int *offsets, int n_models, const float *q, const float *i_exp,
                     const float *sigma, int n_q, const float *scale_range, const float *water_range,
                     const float *expansion_range, int device, double *out, float *out_iq);

/* ---- PDB input (parse_pdb.cpp:35-364; typing tables utils.cpp:201-314) ---- */
/* united-atom form-factor type (0..15) of a PDB atom with label atom_label ("CA") in
 * residue res_name ("GLY"); the element is element_index when >= 0 (a type), else the
 * symbol element ("C", "N", ..., "Zn"; NULL: the label's first letter), unknown symbols
 * typed N.  Replaces both map_pdb_atom_to_ff_type overloads (utils.cpp:292-314). */
int rsx_pdb_atom_type(const char *atom_label, const char *res_name, const char *element, int element_index);
/* ATOM / HETATM records of a PDB file as bodies {x, y, z, type bits}: waters (HOH),
 * hydrogens and alternate locations other than 'A' skipped.  all_models = 0: the first
 * model (parse_pdb<>::first_model_as_complex_atoms); 1: every MODEL, model m is
 * bodies [offsets[m], offsets[m+1]) (all_models_as_complex_atoms).  Call with bodies and
 * offsets NULL for the counts, then with room for n_atoms bodies and n_models + 1
 * offsets.  verbose >= 1 reports elements typed N on stderr.  RSX_PDB_FILE_NOT_FOUND
 * when the file cannot be opened. */
int rsx_pdb_read(const char *path, int all_models, int verbose, float *bodies, long long capacity,
                 long long *n_atoms, int *offsets, int offsets_capacity, int *n_models);

/* ---- batched ensemble (no reference counterpart; calc_saxs/calc_saxs.hpp:611-624 loops frames) ---- */
/* frames: bodies concatenated, frame f is [offsets[f], offsets[f+1]) (offsets[0] = 0, never
 * decreasing); uploads and lays them out.  The batch keeps a layout of its own (q chunks,
 * classes, factor table, jobs), untouched by single-structure calls on the same engine.  With
 * hydration on (water weight != 0, a water row and SASA parameters) each frame's exposed-dot
 * counts are taken once, frame by frame, and its atoms get the classes type x count. */
int rsx_batch_set_frames(rsx_engine *h, const float *bodies, const int *offsets, int n_frames);
/* per-frame I(q) [n_frames][n_q] and the uniform-weight mean [n_q]; either may be NULL.
 * A changed factor table, water weight or expansion factor re-forms only the class factors
 * (the captured launch is kept: a fit over (c, w) pays one factor upload per point); new q,
 * tuning, SASA parameters or hydration switched on lay the kept frames out again. */
int rsx_batch_calc(rsx_engine *h, double *out_frames, double *out_mean);
/* same on resident frames, results left on the device (no host copy); graph-captured */
int rsx_batch_calc_dev_async(rsx_engine *h, double *out_frames_dev, double *out_mean_dev);

/* Batched VJP (no reference counterpart: the reference has no adjoint):
 * grad [atoms of all frames][3] FP32 in the frames' order as given to rsx_batch_set_frames =
 * sum_k w[f][k] dI_f(q_k)/dx_i, f the frame of atom i, for weights w [n_frames][n_q] FP64 (the
 * q = 0 entries are ignored).  One launch of the FP32 adjoint over every frame's pairs and a
 * fixed-order reduction; needs a uniform q grid the recurrence covers.  Synchronous. */
int rsx_batch_vjp(rsx_engine *h, const double *w, int n_frames, int n_q, float *out_grad);
/* batched launch form (FP32 variants): 1 = one persistent grid of resident waves pulling
 * (frame job, q chunk) items from a work counter, longest first (default; taken only where
 * the work-queue kernel keeps the per-item kernel's occupancy); 0 = one wave per item */
int rsx_set_persistent(rsx_engine *h, int on);

/* ---- differentiable fit (no reference counterpart; uses saxs_profile.hpp:137-166 scale/chi) ---- */
/* chi2 = sum_k (I_exp - c I)^2 / sigma^2 at the least-squares c; grad = dchi2/dx [n][3].
 * The chi^2 weights amplify forward errors by I / sigma, so this and rsx_debye_vjp run the FP32
 * per-lane-sum recurrence with runs of at most 51 q (its error grows with the run length): an
 * engine on RSX_VARIANT_FP32R or on longer runs moves there on the first call (rsx_autotune,
 * rsx_set_tuning or rsx_set_variant lift the cap until the next call); an FP64 variant is kept. */
int rsx_chi2_grad(rsx_engine *h, const float *bodies, int n, const float *i_exp, const float *sigma,
                  double *out_chi2, double *out_scale, float *out_grad, double *out_iq);
/* adjoint only: grad_i = sum_k w_k dI_k/dx_i for caller weights w [n_q] (backward of I(q)) */
int rsx_debye_vjp(rsx_engine *h, const float *bodies, int n, const double *w, float *out_grad);
/* device-resident forms of the two above (no host copies; asynchronous on the engine stream,
 * rsx_synchronize reports the bodies' checks as for rsx_calc_curve_dev_async).  They reuse the
 * layout of the last host call (RSX_NO_INITIAL_CALC before one, RSX_ARGS_NOT_SET when the
 * parameters changed since) and need a uniform q grid.
 * vjp: grad_dev [n][3] FP32 = sum_k w_dev[k] dI_k/dx_i, w_dev [n_q] FP64.
 * chi2: i_exp_dev, sigma_dev [n_q] FP32; out_dev = {chi2, scale} FP64; grad_dev [n][3] FP32 (may be
 * NULL); iq_dev [n_q] FP64 (may be NULL).  Needs the gradient configuration of the forward, which
 * a host rsx_chi2_grad sets up (RSX_ARGS_NOT_SET otherwise). */
int rsx_debye_vjp_dev_async(rsx_engine *h, const void *bodies_dev, int n, const double *w_dev, float *grad_dev);
int rsx_chi2_grad_dev_async(rsx_engine *h, const void *bodies_dev, int n, const float *i_exp_dev,
                            const float *sigma_dev, double *out_dev, float *grad_dev, double *iq_dev);
/* adjoint kernel: fp32 1 = FP32 difference-form recurrence (default; its class-pair coefficient
 * form on plain layouts of at most 64 classes, K 64, 100, 128 or 200, the per-lane form otherwise),
 * 2 = the FP32 per-lane form always, 0 = FP64 recurrence; k q values per thread (0 = automatic);
 * col_range columns per job, a multiple of 64 (0 = automatic: 64 or 128).  Applies to
 * rsx_batch_vjp too. */
int rsx_set_adjoint_tuning(rsx_engine *h, int fp32, int k, int col_range);

/* ---- tuning and measurement ---- */
/* choose variant, q-per-thread and column-range length for the current n_q / size; cache in a text file */
int rsx_autotune(rsx_engine *h, const char *cache_path);
/* the same for the batch (no reference counterpart): the FP32 kernels over q per thread, column
 * range and launch form (work queue or one wave per item), timed as graph replays on the
 * frames set; the key is architecture, n_q, frame count and atom total (powers of two), class
 * kind; a cached key is applied without timing */
int rsx_batch_autotune(rsx_engine *h, const char *cache_path);
/* the batch's launch configuration (laid out on demand): variant, q per thread, column range,
 * resident waves of the work-queue form (0: one wave per item); pointers may be NULL */
int rsx_batch_get_tuning(rsx_engine *h, int *variant, int *k, int *col_range, int *persistent_waves);
/* tuning knobs: k (q per thread, 0 = automatic), col_range (atoms per job, 0 = automatic) */
int rsx_set_tuning(rsx_engine *h, int k, int col_range);
/* recurrence kernel for uniform q grids (the next calc_curve lays out again):
 *   RSX_VARIANT_WAVE   FP64 recurrence, one wave per (job, q chunk)
 *   RSX_VARIANT_SHARED FP64 recurrence, the q chunks of a job share the pair geometry through LDS
 *   RSX_VARIANT_FP32   FP32 difference-form recurrence, per-lane sums over class runs
 *   RSX_VARIANT_FP32R  the FP32 recurrence with each column class run reduced over the
 *                      rows when it ends (fewer registers, K up to 128; default) */
#define RSX_VARIANT_WAVE    0
#define RSX_VARIANT_SHARED  1
#define RSX_VARIANT_FP32    2
#define RSX_VARIANT_FP32R   3
int rsx_set_variant(rsx_engine *h, int variant);
/* current q per thread, column range of the forward layout, 1 if the recurrence path is active,
 * and the recurrence variant (any pointer may be NULL) */
int rsx_get_tuning(rsx_engine *h, int *k, int *col_range, int *recurrence, int *variant);
/* launch shape of the current batch graph (no reference counterpart): persistent_waves > 0 when
 * the work-queue kernel runs with that many resident waves, 0 for one wave per (job, q chunk)
 * item; n_jobs = pair jobs over all frames.  RSX_NO_INITIAL_CALC before the first batch launch. */
int rsx_batch_launch(rsx_engine *h, int *persistent_waves, int *n_jobs);
/* the adjoint layout of the last gradient call (batch 0: rsx_chi2_grad / rsx_debye_vjp and their
 * device forms; batch 1: rsx_batch_vjp), no reference counterpart: form 0 = FP64 recurrence,
 * 1 = FP32 class-pair coefficient form, 2 = FP32 per-lane form; k q per thread, col_range columns
 * per job (any pointer may be NULL).  RSX_NO_INITIAL_CALC before such a call. */
#define RSX_ADJOINT_FP64      0
#define RSX_ADJOINT_PAIR      1
#define RSX_ADJOINT_PER_LANE  2
int rsx_get_adjoint_layout(rsx_engine *h, int batch, int *form, int *k, int *col_range);
/* average device time (ms) of the last timed region, measured with HIP events on the engine stream */
int rsx_time_forward(rsx_engine *h, int reps, float *ms_total, float *ms_pair_kernel);
/* average device time (ms) of the adjoint pass (kernel + fixed-order reduction) on the weights of
 * the last rsx_chi2_grad / rsx_debye_vjp call, HIP events on the engine stream */
int rsx_time_adjoint(rsx_engine *h, int reps, float *ms);
/* average device time (ms) of the batched adjoint pass (kernel + reduction) on the weights of
 * the last rsx_batch_vjp call, HIP events on the engine stream (no reference counterpart) */
int rsx_time_batch_vjp(rsx_engine *h, int reps, float *ms);
/* average device time (ms) of one SASA count of the engine's atoms (cell list + dot occlusion
 * kernels), HIP events on the engine stream; needs a prior hydrated calc_curve (no reference
 * counterpart) */
int rsx_time_sasa(rsx_engine *h, int reps, float *ms);
/* plain device-to-device copy of `bytes` (float4 grid-stride kernel), average ms per launch;
 * the known-size reference for HBM counter calibration */
int rsx_time_copy(rsx_engine *h, long long bytes, int reps, float *ms);
/* device memory helpers for callers without another allocator */
int rsx_dev_alloc(rsx_engine *h, void **ptr, long long bytes);
int rsx_dev_free(rsx_engine *h, void *ptr);
int rsx_memcpy_h2d(rsx_engine *h, void *dst, const void *src, long long bytes);
int rsx_memcpy_d2h(rsx_engine *h, void *dst, const void *src, long long bytes);

#ifdef __cplusplus
}
#endif

#endif /* RSX_CAPI_H */
