/*
 * saship.h -- C ABI of libsaship.so, the MI355X (gfx950) I(q) engine.
 *
 * The library evaluates the sasmodels polydispersity-mesh kernel for
 * thirteen form factors and the hard-sphere structure factor on one GPU.  Callers pass the same buffers the reference's
 * compiled kernels take:
 *
 *   void <model>_{Iq|Iqxy|Imagnetic}(int32 nq, int32 pd_start, int32 pd_stop,
 *        const ProblemDetails *details, const double *values, const double *q,
 *        double *result, double cutoff, int32 radius_effective_mode)
 *   (sasmodels/kernel_iq.c:317-328; ctypes argtypes sasmodels/kerneldll.py:369)
 *
 * split into handles: the model (and its quadrature) once, the q vector once
 * per kernel (sasmodels/kernel.py:19-24, make_kernel(q_vectors)), then one
 * call per evaluation (sasmodels/kernel.py:198-205, Kernel._call_kernel).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative code on error;
 *     saship_last_error() then gives a message for the calling thread.
 *   - Buffers are host memory owned by the caller; device memory is owned by
 *     the handles.  Calls are synchronous.
 *   - One HIP stream per kernel/batch handle.  Calls on different handles
 *     may run from different threads; calls on one handle are serialised.
 *   - result keeps the reference layout and accumulate semantics
 *     (kernel_iq.c:392-423, :826-847): for the 1D kernel result[2i] = sum
 *     w*F^2, result[2i+1] = sum w*F; for 2D kernels result[i] = sum w*F^2;
 *     then result[nout*nq + 0..3] = sum w, sum w*V_form, sum w*V_shell,
 *     sum w*R_eff.  pd_start == 0 starts fresh; pd_start > 0 adds to the
 *     caller's result, so a mesh may be split into chunks.
 */
#ifndef SASHIP_H
#define SASHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model ids (parameter tables: sasmodels/models/<name>.py) */
#define SASHIP_SPHERE 0
#define SASHIP_CYLINDER 1
#define SASHIP_ELLIPSOID 2
#define SASHIP_CORE_SHELL_CYLINDER 3
#define SASHIP_CORE_SHELL_SPHERE 4
/* structure factor (sasmodels/models/hardsphere.py): parameters
 * radius_effective, volfraction; nothing polydisperse, so details has the
 * four trailing entries only (max_pd 0); Iq and Iqxy (q = |(qx, qy)|) give
 * result[i] = S(q_i) and sums (1, 1, 1, 0); evaluated in FP64 */
#define SASHIP_HARDSPHERE 5
/* more structure factors with the same conventions (FP64; four parameters
 * radius_effective, volfraction, then perturb, stickiness / welldepth,
 * wellwidth): sasmodels/models/stickyhardsphere.py, squarewell.py */
#define SASHIP_STICKY_HARDSPHERE 14
#define SASHIP_SQUAREWELL 15
/* charged spheres, Hayter-Penfold rescaled MSA (sasmodels/models/
 * hayter_msa.py): six parameters radius_effective, volfraction, charge,
 * temperature, concentration_salt, dielectconst; NaN where the reference's
 * coefficient iteration fails */
#define SASHIP_HAYTER_MSA 16
/* shapes with three orientation angles (theta, phi, psi; xy_mode "qabc"):
 * 1D kernels integrate the orientation with nested rules (adaptive as
 * adaptive.c with outer_n limits, or the fixed size given) */
#define SASHIP_PARALLELEPIPED 6
#define SASHIP_TRIAXIAL_ELLIPSOID 7
/* paracrystals: FP64, Gauss-150 product rule over (theta, phi) unless a
 * size is given; no <F>: the 1D result has one output per q (nq + 4) */
#define SASHIP_BCC_PARACRYSTAL 8
#define SASHIP_FCC_PARACRYSTAL 9
#define SASHIP_SC_PARACRYSTAL 10
/* hollow shapes (result[nout*nq + 2] is the shell volume sum) and the fuzzy sphere */
#define SASHIP_HOLLOW_CYLINDER 11
#define SASHIP_VESICLE 12
#define SASHIP_FUZZY_SPHERE 13
/* concentric shells with vector parameters (sasmodels/models/core_multi_shell.py,
 * onion.py): the control parameter n (0..10) and ten value slots per vector
 * (sld1..sld10, thickness1..thickness10, ...), magnetic triples for every
 * sld slot (modelinfo.py:634-672); onion is evaluated in FP64 */
#define SASHIP_CORE_MULTI_SHELL 17
#define SASHIP_ONION 18
/* more oriented shapes: core_shell_ellipsoid and core_shell_bicelle (theta,
 * phi), elliptical_cylinder (theta, phi, psi; nested 1D rules) */
#define SASHIP_CORE_SHELL_ELLIPSOID 19
#define SASHIP_ELLIPTICAL_CYLINDER 20
#define SASHIP_CORE_SHELL_BICELLE 21
/* models defined by their intensity alone (no <F>: the 1D result has one
 * output per q, nq + 4), evaluated in FP64; 2D uses I(|q|) and the magnetic
 * kernel (models with sld parameters only) re-evaluates I with each spin
 * cross section's effective slds.  Parameter tables: sasmodels/models/<name>.py */
#define SASHIP_GUINIER 22
#define SASHIP_GAUSSIAN_PEAK 23
#define SASHIP_LORENTZ 24
#define SASHIP_DAB 25
#define SASHIP_GEL_FIT 26
#define SASHIP_MASS_FRACTAL 27
#define SASHIP_SURFACE_FRACTAL 28
#define SASHIP_MASS_SURFACE_FRACTAL 29
#define SASHIP_STAR_POLYMER 30
#define SASHIP_MONO_GAUSS_COIL 31
#define SASHIP_FRACTAL 32
#define SASHIP_FRACTAL_CORE_SHELL 33
#define SASHIP_LAMELLAR_HG 34
#define SASHIP_LAMELLAR_STACK_CAILLE 35
#define SASHIP_LAMELLAR_HG_STACK_CAILLE 36
#define SASHIP_LAMELLAR_STACK_PARACRYSTAL 37
#define SASHIP_LINEAR_PEARLS 38
#define SASHIP_POLYMER_MICELLE 39
#define SASHIP_RASPBERRY 40

/* kernel kinds: the three kernels generated per model (generate.py:1144-1180) */
#define SASHIP_KIND_IQ 0        /* <model>_Iq: 1D, <F> and <F^2> */
#define SASHIP_KIND_IQXY 1      /* <model>_Iqxy: 2D */
#define SASHIP_KIND_IMAGNETIC 2 /* <model>_Imagnetic: 2D with spin cross sections */

/* error codes */
#define SASHIP_OK 0
#define SASHIP_EINVAL -1
#define SASHIP_EHIP -2
#define SASHIP_ENOMEM -3

typedef struct saship_model saship_model;
typedef struct saship_kernel saship_kernel;
typedef struct saship_batch saship_batch;

/* Message for the last failed call on this thread ("" if none). */
const char *saship_last_error(void);

/* Number of visible HIP devices.  Replaces the platform probe of
 * sasmodels/core.py:374-450 (parse_dtype) for the "hip" platform. */
int saship_device_count(int32_t *count);

/* Select the device used by handles created afterwards on this thread.
 * Replaces the context set-up of sasmodels/kernelcl.py:246-313. */
int saship_init(int32_t device);

/* Measured FP32 FMA rate of the current device in TFLOP/s (non-packed
 * instructions, eight independent chains per lane, 16 workgroups of 256
 * per CU, `iters` x 64 FMAs per lane): the issue ceiling the FP32
 * evaluation kernels run against.  No reference counterpart. */
int saship_probe_fma(int32_t iters, double *tflops);

/* Model handle: form factor plus orientation quadrature.  quad_n == 0 is
 * the reference default, adaptive 20/76/500/5000 points by q*r_max
 * (models/lib/adaptive.c:11276-11298); quad_n > 0 fixes an n-point
 * Gauss-Legendre rule (generate.py:291-320, set_integration_size).
 * Replaces sasmodels/core.py:293-341 (build_model) for platform "hip". */
int saship_model_create(int32_t model_id, int32_t quad_n, saship_model **out);
int saship_model_release(saship_model *model);

/* Kernel handle: uploads q once.  1D: q[nq]; 2D (is2d != 0): q[2*nq]
 * interleaved (qx, qy), the layout of sasmodels/kernelpy.py:84-94.
 * Replaces KernelModel.make_kernel (sasmodels/kernel.py:19-24). */
int saship_kernel_create(saship_model *model, int32_t nq, const double *q,
                         int32_t is2d, saship_kernel **out);

/* One evaluation over mesh steps [pd_start, pd_stop).  details is the
 * CallDetails int32 buffer (4*max_pd + 4 entries, sasmodels/details.py:73-88),
 * values the packed parameter vector (sasmodels/details.py:225-254).
 * result has nout*nq + 4 doubles, nout = 2 for SASHIP_KIND_IQ of a model with
 * <F>, else 1 (2D kinds, the paracrystals, the structure factor).
 * Replaces Kernel._call_kernel (sasmodels/kernel.py:198-205) as the DLL,
 * OpenCL and CUDA backends implement it (kerneldll.py:449-475,
 * kernelcl.py:599-654, kernelcuda.py:471-518). */
int saship_kernel_call(saship_kernel *kernel, int32_t kind,
                       const int32_t *details, int32_t n_details,
                       const double *values, int32_t n_values,
                       double cutoff, int32_t radius_effective_mode,
                       int32_t pd_start, int32_t pd_stop, double *result);

/* The evaluation of saship_kernel_call without any allocation: no kernel
 * handle, no hipMalloc or hipFree, no allocating library call.  No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  q (nq doubles for SASHIP_KIND_IQ, 2*nq interleaved (qx, qy)
 * otherwise) and result (nout*nq + 4 doubles, the layout and accumulate
 * semantics of saship_kernel_call) are DEVICE memory; details and values are
 * host memory, read before the call returns.  Work is enqueued on `stream`
 * (a hipStream_t; NULL is the null stream) and the call returns without
 * waiting for it.  The intermediate buffers (staged values, FP32 q, mesh
 * records, partial sums) live in one fixed 32 MiB scratch per device,
 * owned by the library until the process ends; calls that do not fit it
 * (mesh records plus q-dependent partial sums) fail with SASHIP_EINVAL.
 * Every call on a device shares that device's scratch: calls whose work
 * overlaps in time on one device race, so serialise them (one stream, or
 * synchronise between calls).  Same evaluation path as saship_kernel_call
 * with the separable cylinder path allowed, except that a cylinder under the
 * adaptive orientation rule runs on the generic mesh kernel (the separable
 * form of that rule works in a batch's buffers). */
int saship_kernel_call_noalloc(saship_model *model, int32_t kind, int32_t nq, const double *q,
                               const int32_t *details, int32_t n_details,
                               const double *values, int32_t n_values,
                               double cutoff, int32_t radius_effective_mode,
                               int32_t pd_start, int32_t pd_stop, double *result, void *stream);

/* Diagnostic, no reference counterpart: one parameter set of structure
 * factor s_model (SASHIP_HARDSPHERE, SASHIP_STICKY_HARDSPHERE,
 * SASHIP_SQUAREWELL, SASHIP_HAYTER_MSA; pars in the order of its kernel
 * call, unused trailing entries ignored): the q-independent coefficients
 * the batched product path solves once per set (prep, 8 doubles: for the
 * Hayter MSA a, b, c, f, eta, gek, k, s with a = NaN where the coefficient
 * iteration fails; zeros for the closed forms) and S at nq q points (sq),
 * computed on the host (on_device == 0) or on the current device by the
 * same code.  q, prep and sq are host memory. */
int saship_structure_eval(int32_t s_model, const double *pars, int32_t nq, const double *q,
                          int32_t on_device, double *prep, double *sq);

/* Evaluation strategy: with separable != 0 (the default) a cylinder 1D call
 * with a mesh over radius and length only is evaluated as bilinear forms of
 * the radius and length factors (same sums, far fewer special-function
 * evaluations): with a fixed Gauss rule directly, with the adaptive rule
 * (quad_n = 0) for a whole unpruned mesh on equally spaced grids through a
 * one-set batch the kernel handle keeps; 0 forces the generic mesh kernel.
 * No reference counterpart. */
int saship_kernel_set_path(saship_kernel *kernel, int32_t separable);

/* Graph replay (default off): the upload and kernels of a call
 * are captured once as a hipGraph and replayed for later calls of the same
 * shape (kind, path, mesh range, cutoff, buffer sizes); enable = 0 launches
 * every call eagerly (measured faster for single calls, whose eager
 * sequence is a few short kernels).  Timed calls are always eager.  No
 * reference counterpart. */
int saship_kernel_set_graph(saship_kernel *kernel, int32_t enable);

/* Kernel timing: when enabled, the device time of the evaluation kernel of
 * every call is summed (HIP events on the handle's stream).  No reference
 * counterpart; used by the benchmark. */
int saship_kernel_set_timing(saship_kernel *kernel, int32_t enable);
int saship_kernel_get_timing(saship_kernel *kernel, double *eval_ms, int32_t *ncalls);

/* Releases the device buffers and stream.  Replaces Kernel.release
 * (sasmodels/kernel.py:190-196). */
int saship_kernel_release(saship_kernel *kernel);

/*
 * Batched evaluation: many parameter sets, each with its own q vector and
 * mesh, in one persistent launch.  No reference counterpart: the reference
 * evaluates a fit population (a bumps DREAM generation, a batch fit) one
 * Kernel call per set (sasmodels/bumps_model.py:250-285,
 * sasmodels/direct_model.py:319-351).  For callers with many independent
 * curves per step; each set's result equals its single-call result.
 */

/* nq[s] q points per set; q concatenated (1D) or (qx, qy) interleaved (2D). */
int saship_batch_create(saship_model *model, int32_t nsets, const int32_t *nq,
                        const double *q, int32_t is2d, saship_batch **out);

/* Whole-mesh evaluation of every set.  details: nsets * n_details int32
 * (one CallDetails buffer per set); values: concatenated value vectors with
 * set s at values[values_offset[s] .. values_offset[s+1]).  Outputs (either
 * may be NULL): result, the raw accumulators of every set back to back in
 * the layout of saship_kernel_call (nout*nq[s] + 4 each); iq, I(q) of every
 * set back to back (sum of nq[s]), normalised as Kernel.Iq does
 * (sasmodels/kernel.py:80-188). */
int saship_batch_call(saship_batch *batch, int32_t kind,
                      const int32_t *details, int32_t n_details,
                      const double *values, const int64_t *values_offset,
                      double cutoff, int32_t radius_effective_mode,
                      double *result, double *iq);

/* Pinned (page-locked, device-mapped, coherent) host memory of `bytes`
 * bytes, and its release.  An `iq` array of saship_batch_call in such
 * memory is written by the call's last kernel directly: no download, no
 * staging buffer and no host copy (that call is launched eagerly, not from
 * the batch's graph, since the array's address changes between calls).  No
 * counterpart in the reference (its OpenCL backend maps its own buffers,
 * sasmodels/kernelcl.py:514-523); for callers that reuse output arrays
 * across calls. */
int saship_host_alloc(int64_t bytes, void **ptr);
int saship_host_free(void *ptr);

/* Resolution smearing of the batch's I(q) (Pinhole1D / Slit1D weight
 * matrices, sasmodels/resolution.py:172-221; Pinhole2D bin averages,
 * resolution2d.py:91-175): per set s, n_out[s] smeared points; for the
 * o-th smeared point of all sets (set order), the weights
 * weights[offset[o] .. offset[o] + length[o]) apply to the set's q points
 * start[o] + k*stride[o], k < length[o] (the set's q are its q_calc;
 * stride == NULL means 1).  Later saship_batch_call results return
 * sum(n_out) smeared I values in `iq` (the raw result stays unsmeared).
 * n_out == NULL switches smearing off.  Every band is validated against
 * the set's q count and n_weights. */
int saship_batch_set_resolution(saship_batch *batch, const int32_t *n_out, const int32_t *start,
                                const int32_t *length, const int64_t *stride, const int64_t *offset,
                                const double *weights, int64_t n_weights);

/* Resolution smearing with repeated bands: as saship_batch_set_resolution,
 * and the o-th smeared point sums its band repeat[o] times, the j-th copy
 * starting repeat_step[o]*j points further on, with the copies weighted by
 * repeat_weights[repeat_offset[o] + j].  This is the 2D slit of an oriented
 * sample (Slit2D, sasmodels/resolution2d.py:177-241): the model on a
 * (qy, qx) grid, the qx band of each point's pinhole weights repeated over
 * the qy rows with the trapezoid weights.  Every band is validated against
 * the set's q count and both weight arrays. */
int saship_batch_set_resolution_repeat(saship_batch *batch, const int32_t *n_out, const int32_t *start,
                                       const int32_t *length, const int64_t *stride, const int64_t *offset,
                                       const double *weights, int64_t n_weights, const int32_t *repeat,
                                       const int64_t *repeat_step, const int64_t *repeat_offset,
                                       const double *repeat_weights, int64_t n_repeat_weights);

/* P@S products on the device: the batch's form factor times structure
 * factor s_model (SASHIP_HARDSPHERE, SASHIP_STICKY_HARDSPHERE,
 * SASHIP_SQUAREWELL or SASHIP_HAYTER_MSA; 0 switches it off), combined in
 * the final kernel as ProductKernel.Iq does (sasmodels/product.py:425-528).
 * pars holds SASHIP_SPARS doubles per set: (radius_effective, volfraction,
 * structure_factor_mode, radius_effective_mode, volfraction_in_form, s2,
 * s3, s4, s5) with s2..s5 the structure factor's further parameters in
 * order, 0 past its last one (all 0 for the hard sphere).
 * Later calls return in `iq`
 *     n * (<F^2> S)                 structure_factor_mode 0
 *     n * (<F^2> + <F>^2 (S - 1))   structure_factor_mode > 0 (beta, 1D only)
 *   + background,
 * with number density n = scale*volfraction/V_shell, or scale/V_shell when
 * volfraction_in_form is non-zero (a form factor with its own volfraction,
 * already in its F); S at the form's mean effective radius
 * (radius_effective_mode > 0; it must equal the call's
 * radius_effective_mode) or at radius_effective (mode 0), and at volume
 * fraction volfraction * V_form/V_shell.  scale and background are the
 * set's values[0..1]; the raw result stays P's. */
#define SASHIP_SPARS 9
int saship_batch_set_structure(saship_batch *batch, int32_t s_model, const double *pars, int32_t nsets);

/* Runs the device work of the last successful saship_batch_call again
 * nsteps times back to back with its inputs already in device memory and
 * its outputs left there (no staging upload, no result download), and
 * returns the elapsed device time of all nsteps in device_ms: the
 * device-resident throughput of the call.  No reference counterpart. */
int saship_batch_replay(saship_batch *batch, int32_t nsteps, double *device_ms);

/* Launch shape: use_graph replays a captured hipGraph for repeated calls of
 * the same shape; pts_per_item is the mesh chunk of one work item (on the
 * separable cylinder path: Gauss nodes per item); waves_per_cu sizes the
 * persistent grid. */
int saship_batch_set_options(saship_batch *batch, int32_t use_graph,
                             int32_t pts_per_item, int32_t waves_per_cu);
int saship_batch_get_options(saship_batch *batch, int32_t *use_graph,
                             int32_t *pts_per_item, int32_t *waves_per_cu);

/* Evaluation strategy of the batch, as saship_kernel_set_path: the
 * separable cylinder path is taken when every set of a call allows it.
 * No reference counterpart. */
int saship_batch_set_path(saship_batch *batch, int32_t separable);

/* Times the device work of a call (staging copy, expansion, evaluation,
 * final sums) over a grid of (pts_per_item, waves_per_cu) on these inputs,
 * keeps the fastest and returns it. */
int saship_batch_autotune(saship_batch *batch, int32_t kind,
                          const int32_t *details, int32_t n_details,
                          const double *values, const int64_t *values_offset,
                          double cutoff, int32_t *best_pts_per_item,
                          int32_t *best_waves_per_cu);

int saship_batch_set_timing(saship_batch *batch, int32_t enable);
int saship_batch_get_timing(saship_batch *batch, double *eval_ms, int32_t *ncalls);
int saship_batch_release(saship_batch *batch);

/*
 * Learned surrogate: MLP 6 -> 4x256 SiLU -> 1, trained and served with BF16
 * MFMA (FP32 accumulate).  No reference counterpart: for callers that need
 * many approximate curves of the polydisperse cylinder family quickly, with
 * the plain kernels as fallback and judge.  All pointers are device memory;
 * `stream` is a hipStream_t (NULL: the default stream).
 *
 * theta: the flat FP32 parameter vector (n_params floats: W1[6][256], b1,
 * W2[256][256], b2, W3, b3, W4, b4, w5[256], b5); pack: n_pack BF16 values
 * refreshed by saship_mlp_pack from theta; work: work_bytes for `batch` rows.
 */
int saship_mlp_sizes(int32_t *n_params, int32_t *n_pack, int64_t *work_bytes, int32_t batch);
int saship_mlp_pack(const float *theta, void *pack, void *stream);
/* y[batch] = network(X[batch][6]) */
int saship_mlp_forward(const float *theta, const void *pack, const float *X, int32_t batch,
                       void *work, float *y, void *stream);
/* Forward, mean squared error against target, and the gradient of the loss
 * with respect to theta into grad (n_params floats); loss_parts receives
 * ceil(batch/256) partial sums of squared residuals (may be NULL). */
int saship_mlp_train_step(const float *theta, const void *pack, const float *X, const float *target,
                          int32_t batch, void *work, float *grad, float *loss_parts, void *stream);
/* Fused inference over ncurves x nq rows: curves[ncurves][5] = radius,
 * length, radius_pd, length_pd, contrast; logq[nq] = log10 q; norm[12] =
 * feature means then standard deviations.  Features of a row: log qR,
 * log qL, radius_pd, length_pd, bL = log(1 + (qL/pi)^2)/2,
 * bR = 3/2 log(1 + (qR)^2/4).  out_iq = 0 writes the network output r;
 * 1 writes I = 1e-4 contrast^2 pi R^2 L exp(r - bL - bR).  grid <= 0 picks
 * the persistent grid size. */
int saship_mlp_infer(const float *theta, const void *pack, const float *curves, int32_t ncurves,
                     const float *logq, int32_t nq, const float *norm, int32_t out_iq, float *y,
                     int32_t grid, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SASHIP_H */
