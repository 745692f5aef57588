/*
 * odtbrain_hip.h -- C ABI of the gfx950 filtered-backpropagation library
 * (libodtbrain_hip.so).
 *
 * The reference (ODTbrain 0.4.12) is pure Python; its hot path is the body
 * of three functions whose arithmetic runs in third-party native code
 * (FFTW via pyfftw, numexpr, scipy.ndimage).  This library replaces that
 * body.  The Python package odtbrain_amd keeps the reference signatures
 * and validation and binds the entry points below with ctypes:
 *
 *   odt_plan_create / odt_plan_execute / odt_plan_destroy
 *     replace the per-call body of
 *       backpropagate_3d        odtbrain/_alg3d_bpp.py:254-555
 *       backpropagate_3d_tilted odtbrain/_alg3d_bppt.py:581-949
 *       backpropagate_2d        odtbrain/_alg2d_bpp.py:142-326
 *     (pad -> FFT2 -> prefactor -> per-z propagator x IFFT2 -> crop ->
 *      spline prefilter -> rotate/affine resample -> sum over angles)
 *   odt_fft_c2c            replaces the pyfftw.FFTW plans
 *                          (_alg3d_bpp.py:460-478) and numpy.fft (_alg2d_bpp.py:264,306)
 *   odt_spline_prefilter   replaces scipy.ndimage.spline_filter(1d) as called
 *                          inside scipy.ndimage.rotate/affine_transform
 *                          (_alg3d_bpp.py:73-81, _alg3d_bppt.py:931-944)
 *   odt_rotate_accumulate  replaces _mprotate/_rotate + ``outarr += arr``
 *                          (_alg3d_bpp.py:35-81, 532-546) on prefiltered planes
 *
 * Conventions: every function returns 0 on success and a nonzero ODT_ERR_*
 * code otherwise; odt_last_error() returns a thread-local message.  No C++
 * exception crosses the ABI.  All data pointers named d_* are device
 * pointers owned by the caller; host pointers are read during the call only.
 * A plan owns its small device tables; the large workspace is passed in by
 * the caller (odt_plan_workspace_size).  A plan is thread-compatible, not
 * thread-safe: use one plan per stream.
 */
#ifndef ODTBRAIN_HIP_H
#define ODTBRAIN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ODT_OK 0
#define ODT_ERR_INVALID 1     /* bad argument (shape, size, null pointer) */
#define ODT_ERR_HIP 2         /* HIP runtime error (message has details) */
#define ODT_ERR_UNSUPPORTED 3 /* valid but not implemented combination */
#define ODT_ERR_WORKSPACE 4   /* workspace too small */

#define ODT_KIND_3D 0         /* backpropagate_3d          (_alg3d_bpp.py:84)  */
#define ODT_KIND_3D_TILTED 1  /* backpropagate_3d_tilted   (_alg3d_bppt.py:374) */
#define ODT_KIND_2D 2         /* backpropagate_2d          (_alg2d_bpp.py:8)   */

#define ODT_FP32 0            /* complex64 data, float32 arithmetic */
#define ODT_FP64 1            /* complex128 data, float64 arithmetic */

#define ODT_PAD_NONE 0        /* no padding          (_alg3d_bpp.py:493-495) */
#define ODT_PAD_EDGE 1        /* np.pad mode="edge"  (_alg3d_bpp.py:496-500) */
#define ODT_PAD_RAMP 2        /* np.pad mode="linear_ramp", end_values=padval
                                 (_alg3d_bpp.py:501-506) */

/* flags */
#define ODT_FLAG_IIR_PREFILTER 1 /* run the separable IIR prefilter kernel on
                                    the slabs instead of folding the prefilter
                                    into the inverse transforms */

/* Plan description.  Geometry is computed by the caller in float64 exactly as
 * the reference computes it (pad sizes _alg3d_bpp.py:285-304; rotation
 * matrices of scipy.ndimage.rotate / affine_transform). */
typedef struct odt_desc {
  int32_t kind;          /* ODT_KIND_* */
  int32_t precision;     /* ODT_FP32 / ODT_FP64 */
  int32_t A;             /* number of projections */
  int32_t Ny, Nx;        /* detector rows/cols (2D: Ny = 1, Nx = N) */
  int32_t lNy, lNx;      /* transform sizes after padding */
  int32_t padyl, padxl;  /* leading pad widths */
  int32_t pad_mode;      /* ODT_PAD_* */
  int32_t intp_order;    /* spline order 0..5 */
  int32_t onlyreal;      /* nonzero: real part only */
  int32_t flags;         /* ODT_FLAG_* */
  int32_t angle_param_stride; /* doubles per angle in angle_params: 6 or 12 */
  double padval_re, padval_im; /* linear-ramp end value */
  double km;             /* 2*pi*nm/res */
  double lD;             /* detector distance */
  double z0;             /* z of slice 0: -ln/2 (3D) or -ln/2+0.5 (2D) */
  double ramp_u, ramp_v; /* ramp |kx*v + ky*u|; u=0, v=1 for the y axis */
  const double *weights; /* host, A entries: angle weights (1 if unweighted) */
  /* host, A*angle_param_stride entries.
   * planar (stride 6): m00 m01 m10 m11 off0 off1, input = M @ (o_z, o_x) + off
   *   (the matrix and offset of scipy.ndimage.rotate, axes (z, x));
   * affine (stride 12): 3x3 row-major Dinv then offset, input index
   *   = Dinv @ (x, Ny-1-y, z) + offset in the (x, -y, z) frame of
   *   _alg3d_bppt.py:904-936. */
  const double *angle_params;
  int64_t slab_budget_bytes; /* upper bound for the batch of filtered slabs
                                (0: library default) */
} odt_desc;

typedef struct odt_plan odt_plan;

/* Build a plan: validates the description, derives the transform support,
 * uploads twiddles, filter and geometry tables.  Allocates device memory for
 * the tables only. */
int odt_plan_create(const odt_desc *desc, odt_plan **plan);

/* Launch-shape choices of a plan that do not change what it computes: the
 * reference has none (it has no device code); they are what the autotuner
 * (odtbrain_amd.tune) searches.  0 in a field: the planner's default.
 * Explicit values take precedence over the ODTBRAIN_AMD_* environment
 * switches.  t_budget_bytes only re-chunks the same kernels' arithmetic
 * (bit-identical results); rotate_group regroups the float32 running sums
 * between launches (float64 plans stay bit-identical); rotate_groups_2d
 * regroups the sums over angles and col_per_angle selects another kernel
 * for the same values (both: rounding-level differences). */
typedef struct odt_tuning {
  int32_t rotate_group;     /* angles per rotate-accumulate launch (3D)   */
  int32_t rotate_groups_2d; /* concurrent angle groups of the 2D rotate    */
  int64_t t_budget_bytes;   /* budget of the intermediate transform chunk  */
  int32_t col_per_angle;    /* 1: column stage with one block per angle,
                               2: propagator held across a chunk's angles  */
  int32_t reserved;
} odt_tuning;

/* odt_plan_create with launch-shape choices (tuning may be null). */
int odt_plan_create_tuned(const odt_desc *desc, const odt_tuning *tuning,
                          odt_plan **plan);

/* Bytes of caller-provided device workspace that odt_plan_execute (and
 * odt_plan_adjoint) needs. */
int odt_plan_workspace_size(const odt_plan *plan, size_t *bytes);

/* Run the full reconstruction on `stream` (a hipStream_t, 0 = null stream).
 * d_usin: (A, Ny, Nx) complex64 (FP32) or complex128 (FP64), C order.
 * d_out:  (Nx, Ny, Nx) complex, or real if onlyreal; overwritten.
 * The call is asynchronous with respect to the host.  With
 * ODTBRAIN_AMD_OVERLAP=1 at plan creation, 3D plans whose angles take two
 * filter chunks or more run the row transforms on a second stream owned by
 * the plan (forked from and joined back into `stream` with events, so the
 * work is ordered on `stream` as a whole); under graph capture the call
 * keeps to `stream`. */
int odt_plan_execute(odt_plan *plan, const void *d_usin, void *d_out,
                     void *d_workspace, size_t workspace_bytes, void *stream);

int odt_plan_destroy(odt_plan *plan);

/* Allocation-free variant of odt_plan_execute.  There is no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The workspace is a static device array of ODT_SCRATCH_BYTES per GPU,
 * held by the library until the process ends; plans whose
 * odt_plan_workspace_size exceeds it are rejected with ODT_ERR_WORKSPACE
 * (create them with a smaller slab_budget_bytes).  The call allocates and
 * frees nothing.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (different streams or threads) race and must
 * be serialised by the caller.  Must be called with the plan's device
 * current; timing must be disabled.  Same results as odt_plan_execute. */
#define ODT_SCRATCH_BYTES (128u << 20)
int odt_plan_execute_noalloc(odt_plan *plan, const void *d_usin, void *d_out,
                             void *stream);

/* Kernel timing of the last execute (milliseconds, HIP events on the plan's
 * stream), for the rotate-accumulate stage and the whole call, and the number
 * of rotate-accumulate kernel launches the stage made.  Timing is recorded
 * only when enabled. */
int odt_plan_enable_timing(odt_plan *plan, int enable);
int odt_plan_last_timing(const odt_plan *plan, double *rotate_ms,
                         double *filter_ms, double *total_ms,
                         int *rotate_launches);

/* Adjoint of odt_plan_execute: d_gusin (A, Ny, Nx) = M^H d_gout, where M is
 * the (complex-linear) map from the sinogram to the volume that execute
 * applies.  There is no counterpart in the reference, which has no
 * gradient; it is the backward pass of a differentiable reconstruction.
 * Supported: ODT_KIND_3D plans with complex output, spline order 2..5
 * (with ODT_FLAG_IIR_PREFILTER, or the fused prefilter of orders 2 and 3,
 * whose transpose the transforms carry), edge or no padding, padded sizes
 * 64..512.  Uses the plan's workspace (same size as execute). */
int odt_plan_adjoint(odt_plan *plan, const void *d_gout, void *d_gusin,
                     void *d_ws, size_t ws_bytes, void *stream);

/* Batches of reconstructions replayed as one HIP graph.  The reference
 * reconstructs one sinogram per call (and per Python process pool); a
 * batch runs many plans of uneven sizes back to back on one stream:
 * odt_batch_create executes every plan once (validation, warm-up), then
 * captures the same executions into a graph; odt_batch_launch replays it.
 * The buffers are fixed at creation: d_usin[i] / d_out[i] as for
 * odt_plan_execute, one workspace shared by all entries (at least the
 * largest odt_plan_workspace_size).  stream must be a created stream (the
 * null stream cannot be captured); plans must not have timing enabled. */
typedef struct odt_batch odt_batch;
int odt_batch_create(int n, odt_plan *const *plans, const void *const *d_usin,
                     void *const *d_out, void *d_ws, size_t ws_bytes,
                     void *stream, odt_batch **batch);
/* The same with the plans spread over `lanes` concurrent branches of the
 * graph (longest first onto the least loaded lane): d_ws holds `lanes`
 * equal parts, each at least the largest odt_plan_workspace_size of the
 * plans; the plans of one lane run in order on that lane's part.  Small
 * reconstructions then overlap on the GPU instead of running one after
 * another.  lanes = 1 is odt_batch_create. */
int odt_batch_create_lanes(int n, odt_plan *const *plans, const void *const *d_usin,
                           void *const *d_out, void *d_ws, size_t ws_bytes, int lanes,
                           void *stream, odt_batch **batch);
int odt_batch_launch(odt_batch *batch, void *stream);
int odt_batch_destroy(odt_batch *batch);

/* Batched 1D complex transforms along the last axis, unnormalised.
 * direction -1: forward (exp(-2 pi i jk/n)), +1: backward.  n arbitrary
 * (power of two: Stockham; other lengths up to 2048: Bluestein's chirp
 * convolution on power-of-two transforms; longer: direct DFT), as the
 * reference's FFTW plans take any length (_alg3d_bpp.py:460-478).  In place
 * allowed. */
int odt_fft_c2c(int precision, int64_t batch, int32_t n, int32_t direction,
                const void *d_in, void *d_out, void *stream);

/* Mirror-boundary B-spline prefilter (order 2..5) of `nlines` lines of length
 * n, element (line l, position i) at d_data[l*line_stride + i*elem_stride]
 * (strides in elements of the real type), in place, computed in float64
 * internally for ODT_FP64 and float32 for ODT_FP32. */
int odt_spline_prefilter(int precision, int order, int64_t nlines,
                         int32_t n, int64_t line_stride, int64_t elem_stride,
                         void *d_data, void *stream);

/* Rotate-accumulate on prefiltered real planes:
 * d_coef: (nang, ny, nz, nx) coefficient planes (plane a,y has shape nz x nx),
 * d_out:  (nz, ny, nx), d_out[z,y,x] += sum_a interp(coef[a,y], M_a @ (z,x) + off_a)
 * with scipy.ndimage's constant-mode rule (outside [0,n-1] contributes 0).
 * params: host, nang*6 doubles (m00 m01 m10 m11 off0 off1); each 2x2
 * matrix must be a rotation (|m00|+|m01| and |m10|+|m11| at most sqrt(2),
 * else ODT_ERR_INVALID): the kernel's LDS windows are sized for rotations.
 * accumulate: 0 overwrites d_out, 1 adds to it. */
int odt_rotate_accumulate(int precision, int order, int32_t nang, int32_t ny,
                          int32_t nz, int32_t nx, const void *d_coef,
                          const double *params, void *d_out, int accumulate,
                          void *stream);

/* Object function to refractive index, elementwise over n values
 * (odtbrain/_translate_ri.py:5-57, odt_to_ri): ri = nm * sqrt(f / km^2 + 1)
 * with km = 2 pi nm / res (km^2 formed in float64 on the host), the
 * principal root (non-negative real part) for complex f.  is_complex: 1 for
 * interleaved complex f and ri, 0 for real (a negative radicand gives NaN,
 * as numpy's real sqrt does).  In place allowed; asynchronous. */
int odt_odt_to_ri(int precision, int is_complex, int64_t n, double res,
                  double nm, const void *d_f, void *d_ri, void *stream);

/* Optical projection tomography variant (_translate_ri.py:60-90,
 * opt_to_ri): ri = nm + f / (2 pi) * res, elementwise as above. */
int odt_opt_to_ri(int precision, int is_complex, int64_t n, double res,
                  double nm, const void *d_f, void *d_ri, void *stream);

/* Sinogram preparation (odtbrain/_prepare_sino.py:7-182).
 * rytov = 1: sinogram_as_rytov, d_out complex (A, [Ny,] Nx) =
 *   u0 * (log|u/u0| + i * unwrapped angle(u/u0));
 * rytov = 0: sinogram_as_radon, d_out real = unwrapped angle(u).
 * ndim 2: (A, Nx) sinogram, numpy.unwrap along the detector axis;
 * ndim 3: (A, Ny, Nx), every angle's image unwrapped with the reliability-
 *   sorted algorithm of skimage.restoration.unwrap_phase (Herraez et al.
 *   2002; border reliabilities fixed, ties in edge creation order).
 * align = 1: align_unwrapped (_prepare_sino.py:7-46).  d_u0: null for
 * u0 = 1, else a complex array of the full shape.  Ny ignored for ndim 2.
 * Workspace from odt_prepare_workspace_size; asynchronous. */
int odt_prepare_workspace_size(int precision, int ndim, int32_t A, int32_t Ny,
                               int32_t Nx, int rytov, size_t *bytes);
int odt_sinogram_prepare(int precision, int ndim, int32_t A, int32_t Ny,
                         int32_t Nx, const void *d_usin, const void *d_u0,
                         int rytov, int align, void *d_out, void *d_ws,
                         size_t ws_bytes, void *stream);

/* Missing-apple-core correction (odtbrain/apple.py:124-262, correct) of an
 * object function f of shape (n0, n1, n2) with n0 == n2 (the y axis n1 is
 * the axis of symmetry): iterative projections with the non-negativity
 * constraint (method 0, apple.py:52-71) or the symmetric-histogram
 * constraint (method 1, apple.py:74-121, with its zero-frequency update
 * apple.py:235-237), computed in complex128.
 * d_f complex (f_precision), d_out complex (out_precision; the reference
 * returns complex64).  d_mask: null or n0*n1*n2 bytes (nonzero = may be
 * set to zero).  bg_shell_width < 0: no background shell.
 * enforce_envelope in [0, 1] (0 disables the envelope).  Runs max_iter
 * iterations at most, stopping as the reference does; the number run is
 * written to *iterations.  Synchronises the stream once per iteration.
 * Workspace from odt_apple_workspace_size. */
int odt_apple_workspace_size(int32_t n0, int32_t n1, int32_t n2, size_t *bytes);
/* The apple core mask alone (apple.py:7-49, apple_core_3d): d_core gets
 * n0*n1*n2 bytes, 1 inside the core; n0 == n2 required. */
int odt_apple_core(int32_t n0, int32_t n1, int32_t n2, double res, double nm,
                   uint8_t *d_core, void *stream);
int odt_apple_correct(int f_precision, int32_t n0, int32_t n1, int32_t n2,
                      const void *d_f, double res, double nm, int method,
                      const uint8_t *d_mask, double bg_shell_width,
                      double enforce_envelope, int32_t max_iter,
                      double min_diff, int out_precision, void *d_out,
                      void *d_ws, size_t ws_bytes, void *stream,
                      int32_t *iterations);

/* Envelope of the Fourier amplitudes outside the apple core
 * (odtbrain/apple.py:288-353, envelope_gauss): |F| with the core filled
 * from the (1,0,2) transpose and, where that leaves gaps, from the nearest
 * valid sample, Gaussian-filtered (sigma = max(shape)/100, truncate 4,
 * zero outside).  d_ftdata complex128, d_core n0*n1*n2 bytes, d_env
 * float64; n0 == n2; workspace from odt_apple_workspace_size;
 * asynchronous. */
int odt_apple_envelope(int32_t n0, int32_t n1, int32_t n2, const void *d_ftdata,
                       const uint8_t *d_core, void *d_env, void *d_ws, size_t ws_bytes,
                       void *stream);

const char *odt_last_error(void);
int odt_version(void);

/* Device allocations the library holds (plans' tables and the temporary
 * buffers of running calls).  No counterpart in the reference; it is for
 * tests of the error paths: with the environment variable
 * ODTBRAIN_AMD_FAULT="alloc:N" the N-th device allocation after the variable
 * took that value fails with hipErrorOutOfMemory, with "launch:N" the N-th
 * checked launch or HIP call reports hipErrorLaunchFailure after it was
 * issued; each fault fires once per value.  A failed call must leave this
 * count where it was. */
int odt_debug_live_allocations(int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* ODTBRAIN_HIP_H */
