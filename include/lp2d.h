/*
 * lp2d.h - C ABI of the MI355X (gfx950) angular-spectrum propagation library.
 *
 * This is the seam that replaces lightprop2d's FFT / elementwise backend
 * (`xp` module + `_fft2` / `_ifft2`, lightprop2d/beam2d.py:105-109, 207-229) and the
 * operator bodies built on it (propagate :462-490, lens :492-511,
 * coordinate_filter :310-342, spectral_filter :344-373, iprofile :706-721).
 * The Python `Beam2D` in lightprop2d_amd/beam2d.py calls these entry points.
 *
 * Conventions
 *  - Every data pointer is a DEVICE pointer owned by the caller.  Fields are
 *    row-major (batch, n, n) arrays, rows = y, of complex64 (float re, im) or
 *    complex128 (double re, im), selected by `dtype`.
 *  - n must be a power of two, 2 <= n <= lp2d_max_n().
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  Calls are
 *    stream-ordered and asynchronous.  Device memory is allocated only by lp2d_init
 *    (twiddle tables, once per device; lp2d_plan_create runs it) and
 *    lp2d_batch_create (work tables); workspaces are the caller's.
 *  - Every call returns LP2D_OK (0) or a negative status; lp2d_last_error() gives
 *    the message for the calling thread.  No C++ exception crosses the ABI.
 *  - Forward FFT is unnormalised, inverse scales by 1/n^2 (numpy convention,
 *    beam2d.py:207-229).
 *  - Units are the reference's: centimetres (beam2d.py:37-41); k0 = 1/wl.
 */
#ifndef LP2D_H
#define LP2D_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LP2D_OK 0
#define LP2D_E_ARG (-1)         /* bad argument (null pointer, bad flag)           */
#define LP2D_E_SIZE (-2)        /* n not a supported power of two, bad batch        */
#define LP2D_E_HIP (-3)         /* a HIP runtime call failed                         */
#define LP2D_E_DTYPE (-4)       /* unknown field or mask dtype                       */
#define LP2D_E_INIT (-5)        /* library not initialised on the current device     */

/* field dtypes */
#define LP2D_C64 0
#define LP2D_C128 1

/* mask / multiplier dtypes (lp2d_mask.dtype) */
#define LP2D_MASK_U8 1          /* bool / uint8                */
#define LP2D_MASK_F32 2
#define LP2D_MASK_F64 3
#define LP2D_MASK_C64 4
#define LP2D_MASK_C128 5

/* evanescent-mode rule (lp2d_optics.evanescent) */
#define LP2D_EVANESCENT_PASS 0  /* reference: Kz := 0 so H = 1 (beam2d.py:307-308) */
#define LP2D_EVANESCENT_ZERO 1  /* opt-in: H = 0                                      */
#define LP2D_EVANESCENT_DECAY 2 /* opt-in: H = exp(-2*pi*sqrt(f^2 - k0^2)*|z|)        */

/* An elementwise multiplier broadcast over (batch, y, x) with element strides
 * (0 = broadcast along that axis).  Matches `field *= mask` with numpy broadcasting
 * of the reference's f_gen(X, Y) / f_init arrays (beam2d.py:338, 341, 369, 372). */
typedef struct lp2d_mask {
  const void* ptr;
  int32_t dtype;
  int32_t reserved;
  int64_t stride_b, stride_y, stride_x;
} lp2d_mask;

/* Optical parameters of one call. */
typedef struct lp2d_optics {
  double area_size;       /* cm, grid extent; dL = area_size / n          */
  double wl;              /* cm, wavelength; k0 = 1/wl                    */
  double z;               /* cm, propagation distance (0 = no H)          */
  double aperture_radius; /* cycles/cm; > 0 keeps Kx^2+Ky^2 <= radius^2   */
  double lens_f;          /* cm; != 0 applies the reference lens phase     */
  int32_t evanescent;     /* LP2D_EVANESCENT_*                             */
  int32_t reserved;
} lp2d_optics;

/* ---------------------------------------------------------------- library */
/* Prepares the twiddle tables on `device` (idempotent, thread-safe; allocates).
 * Every other entry point calls it for the current device on first use; a first use
 * on a stream that is being captured into a graph returns LP2D_E_INIT instead, so
 * callers that capture call lp2d_init beforehand. */
int lp2d_init(int device);
const char* lp2d_last_error(void);
int lp2d_version(void);
int lp2d_max_n(int dtype);

/* ---------------------------------------------------------------- grids */
/* Kz grid, FP64 (n, n): nan_to_num(sqrt(k0^2 - Kx^2 - Ky^2)) with Kx = fftfreq(n, dL).
 * Replaces beam2d.py:299-308 (the kernels regenerate it on chip; this is for the
 * `Kz` attribute). */
int lp2d_kz_grid(int n, const lp2d_optics* optics, double* kz_out, void* stream);

/* ---------------------------------------------------------------- Beam2D state ops */
/* out = fft2(in) (inverse = 0) or ifft2(in) (inverse = 1); in == out allowed.
 * Replaces _fft2 / _ifft2 (beam2d.py:207-229). */
int lp2d_fft2(int n, int batch, int dtype, const void* in, void* out, int inverse,
              void* stream);

/* field <- field * masks[0] * ... * lens(optics->lens_f) * exp(i*phase);
 * spectrum_out <- fft2(field) when spectrum_out != NULL.
 * Replaces coordinate_filter (beam2d.py:335-342) and lens (beam2d.py:508-511).
 * nmasks <= 2; phase (float32 radians, (n, n) or (batch, n, n) via phase_stride_b)
 * may be NULL; optics may be NULL when no lens is applied. */
int lp2d_spatial_apply(int n, int batch, int dtype, void* field, const lp2d_mask* masks,
                       int nmasks, const lp2d_optics* optics, const float* phase,
                       int64_t phase_stride_b, void* spectrum_out, void* stream);

/* spectrum <- spectrum * H(optics->z) * aperture * masks[0..nmasks);
 * field_out <- ifft2(spectrum).  With z != 0 this is propagate (beam2d.py:481-490);
 * with masks it is spectral_filter (beam2d.py:366-373). */
int lp2d_spectral_apply(int n, int batch, int dtype, void* spectrum, const lp2d_mask* masks,
                        int nmasks, const lp2d_optics* optics, void* field_out, void* stream);

/* |field|^2 into float (C64) or double (C128) (beam2d.py:717). */
int lp2d_intensity(int64_t count, int dtype, const void* field, void* out, void* stream);

/* Intensity moments for the beam diagnostics (centroid, D4sigma; beam2d.py:669-704).
 * Weights I = |field|^2 (computed in the field's precision, accumulated in FP64),
 * coordinates X = (ix - n/2)*dL, Y = (iy - n/2)*dL.  For each field b, out[8*b + k]:
 *   k=0 sum I, 1 Xc, 2 Yc, 3 sigma_x, 4 sigma_y, 5..7 scratch
 * with Xc = sum(I X)/sum I and sigma_x = sqrt(sum(I (X - Xc)^2)/sum I) (two passes).
 * out is a device array of 8*batch doubles; it is cleared by the call. */
int lp2d_moments(int n, int batch, int dtype, const void* field, double dL, double* out,
                 void* stream);

/* ---------------------------------------------------------------- generators */
/* Field / mask generators of lightprop2d/profiles.py evaluated on the device over the
 * broadcast grid (y[j], x[i]), out[j*nx + i] (the reference calls them as
 * f(X, Y) with X of shape (n,) and Y of shape (n, 1), beam2d.py:302, 338).
 *   LP2D_PROF_PLANE        1                                 (profiles.py:35-38)
 *   LP2D_PROF_GAUSSIAN     a0*exp(-((x-x0)^2 + (y-y0)^2)/rho0^2)       (:64-71)
 *   LP2D_PROF_ROUND_HOLE   exp(-((x-x0)^2+(y-y0)^2)/r^2) >= thresh     (:74-81)
 *   LP2D_PROF_RECT_HOLE    |x-x0| < dx/2 and |y-y0| < dy/2             (:116-128)
 *   LP2D_PROF_SLITS        square holes of side dx at x0 -+ slits/2    (:131-136)
 *   LP2D_PROF_RANDOM       kron(rand, ones(b, b))                      (:29-61)
 *   LP2D_PROF_RANDOM_HOLE  kron(rand, ones(b, b)) * round hole         (:84-101)
 *   LP2D_PROF_RANDOM_PHASE round hole * exp(2*pi*i * kron(rand) * hole) (:104-113)
 * The random kinds take `rand`, a device array of (ny/b) x (nx/b) doubles drawn on
 * the host (the reference's numpy RNG stream); the others ignore it.  thresh is the
 * caller's 1/exp(0.5) so the comparison uses the reference's own constant.
 * out_dtype: LP2D_MASK_U8 (bool), LP2D_MASK_F64 or LP2D_MASK_C128. */
#define LP2D_PROF_PLANE 0
#define LP2D_PROF_GAUSSIAN 1
#define LP2D_PROF_ROUND_HOLE 2
#define LP2D_PROF_RECT_HOLE 3
#define LP2D_PROF_SLITS 4
#define LP2D_PROF_RANDOM 5
#define LP2D_PROF_RANDOM_HOLE 6
#define LP2D_PROF_RANDOM_PHASE 7
typedef struct lp2d_profile_params {
  double a0, rho0, r, thresh;
  double x0, y0, dx, dy, slits;
  int32_t binning;
  int32_t reserved;
} lp2d_profile_params;
int lp2d_profile(int kind, int nx, int ny, const double* x, const double* y,
                 const lp2d_profile_params* params, const double* rand, void* out,
                 int out_dtype, void* stream);

/* ---------------------------------------------------------------- resampling */
/* Beam2D.expand (beam2d.py:375-404): (batch, n, n) -> (batch, n_out, n_out) with the
 * field's edge rows/columns replicated (n_out - n even, d = (n_out - n)/2 on each
 * side); in and out must not overlap. */
int lp2d_expand(int n, int n_out, int batch, int dtype, const void* in, void* out,
                void* stream);
/* Beam2D.coarse (beam2d.py:437-460): block mean over mean_order x mean_order,
 * (batch, n, n) -> (batch, n/mean_order, n/mean_order); mean_order must divide n. */
int lp2d_coarse(int n, int mean_order, int batch, int dtype, const void* in, void* out,
                void* stream);

/* Bilinear resampling on a regular grid: scipy's interp2d(x, x, field, kind='linear')
 * (scipy < 1.14) evaluated at ascending queries, the call Beam2D.crop (beam2d.py:406-435)
 * and Beam2D.lens_image (beam2d.py:513-536) make.  It is FITPACK's degree-1 spline
 * with a breakpoint at every sample: a query outside [t_0, t_{len-1}] is clamped to
 * the edge, the interval is the last l <= len-2 with t_l <= q, and the weights are
 * (t_{l+1} - q)/(t_{l+1} - t_l) and (q - t_l)/(t_{l+1} - t_l), all in FP64.
 * Both axes use the same coordinates.  Source: the win_len x win_len window at
 * (win_off, win_off) of each n x n field; window sample k sits at
 * t_k = (k + src_i0) * src_dx.  Output sample j is taken at
 * q_j = ((dst_sign * j + dst_i0) * dst_dx) * dst_scale.  The caller orders the queries
 * as interp2d sorts them; dst_sign = -1 covers a negative dst_scale.
 * out (batch, n_out, n_out) = scale * interp * exp(-i * phase_coef * (x^2 + y^2)),
 * with x_j = (j + grid_i0) * grid_dl (phase_coef = 0: no chirp).  in and out must not
 * overlap. */
typedef struct lp2d_interp_desc {
  int32_t win_off, win_len; /* window of the source field, both axes (win_len >= 2) */
  int32_t src_i0;           /* t_k = (k + src_i0) * src_dx */
  int32_t dst_i0, dst_sign; /* q_j = ((dst_sign * j + dst_i0) * dst_dx) * dst_scale */
  int32_t grid_i0;          /* chirp coordinate x_j = (j + grid_i0) * grid_dl */
  double src_dx;            /* > 0 */
  double dst_dx, dst_scale;
  double scale;             /* output factor */
  double phase_coef;        /* chirp coefficient [rad / cm^2] */
  double grid_dl;
} lp2d_interp_desc;
int lp2d_interp_linear(int n, int n_out, int batch, int dtype, const void* in, void* out,
                       const lp2d_interp_desc* desc, void* stream);

/* ---------------------------------------------------------------- mode bases */
/* Mode-basis projection (Beam2D.deconstruct_by_modes 'scalar', beam2d.py:563-603):
 * coeffs[b, k] = sum_p modes[k, p] * field[b, window(p)] -- the reference's
 * unconjugated product of the (nmodes, nb*nb) mode matrix with the central nb x nb
 * window (offset (n - nb)/2) of each n x n field, flattened row-major.  modes and field
 * in `dtype`; coeffs is (batch, nmodes) complex128 (FP64 accumulation), overwritten. */
int lp2d_modes_project(int n, int nb, int nmodes, int batch, int dtype, const void* modes,
                       const void* field, void* coeffs, void* stream);
/* The same projection on the matrix cores (FP64 MFMA, every product and sum in FP64)
 * for batches of fields: each workgroup reduces a pixel range of a 64-mode x 32-field
 * tile into `work`, and a second kernel sums the ranges into coeffs.  work_bytes must
 * be at least lp2d_modes_project_ws_size(n, nb, nmodes, batch); work must not alias the
 * other arrays, and its contents are undefined on return.  work == NULL runs
 * lp2d_modes_project.  Same reference interface as lp2d_modes_project
 * (beam2d.py:563-603). */
long long lp2d_modes_project_ws_size(int n, int nb, int nmodes, int batch);
int lp2d_modes_project_ws(int n, int nb, int nmodes, int batch, int dtype, const void* modes,
                          const void* field, void* coeffs, void* work, long long work_bytes,
                          void* stream);
/* Reconstruction (Beam2D.construct_by_modes, beam2d.py:653-667): field[b] =
 * sum_k coeffs[b, k] * mode_k zero-padded into the centre of the n x n grid.  coeffs
 * (batch, nmodes) complex128; field (batch, n, n) in `dtype`, overwritten. */
int lp2d_modes_construct(int n, int nb, int nmodes, int batch, int dtype, const void* modes,
                         const void* coeffs, void* field, void* stream);

/* ---------------------------------------------------------------- field -> field */
/* field_out = ifft2(H(z) * aperture * fft2(field_in * exp(i*phase) * lens)), fused;
 * spectrum_out (may be NULL) receives H*aperture*fft2(...), i.e. the reference's
 * `spectrum` state after the same calls.  field_in == field_out allowed.
 * Reference sequence: coordinate_filter(f_init=exp(i*phase)) -> lens -> propagate(z)
 * -> spectral_filter(aperture) (beam2d.py:339-342, 508-511, 481-490, 370-373). */
int lp2d_propagate(int n, int batch, int dtype, const void* field_in, void* field_out,
                   const float* phase, int64_t phase_stride_b, const lp2d_optics* optics,
                   void* spectrum_out, void* stream);

/* Multi-plane stack: u_0 = field_in, u_{k+1} = the lp2d_propagate of u_k with phase
 * screens[k] (optics->lens_f applied on every plane), k = 0 .. nscreens-1, and
 * field_out = u_nscreens.  screens: nscreens float32 (n, n) planes (radians, shared
 * by the batch), plane k at screens + k*n*n.  planes_out: NULL, or nscreens host-side
 * pointers of which planes_out[k] (1 <= k < nscreens), when not NULL, receives u_k
 * (batch, n, n; must not alias the fields); planes_out[0] is ignored.  field_in ==
 * field_out allowed.  For complex64 sizes with a block-tiled intermediate the inverse
 * row pass of plane k, the screen of plane k+1 and its forward row pass run as one
 * sweep (2*nscreens + 1 sweeps in all).  Reference sequence per plane:
 * coordinate_filter(f_init=exp(i*phi_k)) then propagate(z) (beam2d.py:339-342,
 * 481-490). */
int lp2d_propagate_stack(int n, int batch, int dtype, const void* field_in, void* field_out,
                         const float* screens, int nscreens, const lp2d_optics* optics,
                         void* const* planes_out, void* stream);

/* Adjoint of lp2d_propagate_stack (same n, batch, screens and optics): from grad_out
 * = dL/d(field_out) to grad_in = dL/d(field_in) and, when grad_screens is not NULL,
 * the batch-summed screen gradients (nscreens float32 (n, n) planes, plane k at
 * grad_screens + k*n*n), with lp2d_propagate_adjoint's conventions.  planes_in:
 * nscreens host-side pointers, planes_in[k] = u_k, the input of plane k
 * (planes_in[0] = field_in; the others as lp2d_propagate_stack's planes_out gives
 * them); none may alias grad_in.  grad_in is the passes' intermediate plane (its
 * contents undefined with LP2D_ADJ_NO_GRAD_IN) and may alias grad_out.  work: device
 * memory of work_bytes >= lp2d_propagate_stack_adjoint_ws_size(n, batch, dtype)
 * (LP2D_E_SIZE otherwise; NULL allowed when that is 0 or grad_screens is NULL), not
 * aliasing the other arrays, contents undefined on return.  For complex64 sizes with
 * a block-tiled intermediate the inverse row pass of plane k, the fold and gradient
 * of screen k and the forward row pass of plane k-1's adjoint run as one sweep
 * (2*nscreens + 1 sweeps in all), else one lp2d_propagate_adjoint per plane.  No
 * counterpart in the reference (it has no gradients). */
int lp2d_propagate_stack_adjoint(int n, int batch, int dtype, const void* grad_out,
                                 const void* const* planes_in, const float* screens,
                                 int nscreens, const lp2d_optics* optics, void* grad_in,
                                 float* grad_screens, int adj_flags, void* work,
                                 int64_t work_bytes, void* stream);
/* Bytes of work lp2d_propagate_stack_adjoint needs for (n, batch, dtype) (0: none;
 * -1: invalid arguments). */
int64_t lp2d_propagate_stack_adjoint_ws_size(int n, int batch, int dtype);

/* Adjoint (vector-Jacobian product) of lp2d_propagate, for a real loss with
 * PyTorch's complex-gradient convention (grad = dL/dRe + i dL/dIm):
 *   g1 = ifft2(conj(H) * aperture * fft2(grad_out))
 *   grad_in = conj(exp(i*phase) * lens) * g1
 *   gphi[b] = Im(conj(field_in[b] * exp(i*phase) * lens) * g1[b])
 * grad_phase (float32, (batch, n, n)) may be NULL, and must be NULL when phase is.
 * With a per-field phase (phase_stride_b = n*n) plane b receives gphi[b]; with a
 * shared screen (phase_stride_b = 0) plane 0 receives sum_b gphi[b] and the other
 * planes are scratch.  grad_in is always required (the passes use it as their
 * intermediate plane); with LP2D_ADJ_NO_GRAD_IN in adj_flags its contents are
 * undefined on return (for callers that only need the phase gradient).
 * No counterpart in the reference (it has no gradients). */
#define LP2D_ADJ_NO_GRAD_IN 1
int lp2d_propagate_adjoint(int n, int batch, int dtype, const void* grad_out,
                           const void* field_in, const float* phase, int64_t phase_stride_b,
                           const lp2d_optics* optics, void* grad_in, float* grad_phase,
                           int adj_flags, void* stream);
/* lp2d_propagate_adjoint with a workspace: `work` (batch x n x n of `dtype`, or NULL
 * for the call above) receives the column pass's output, so the last pass reads one
 * buffer and writes another.  The intermediate plane may then stay block-tiled for
 * every gradient schedule (in place, the gradient pass would have to convert whole
 * 8-row blocks).  work must not alias grad_out, field_in or grad_in; its contents are
 * undefined on return.  No counterpart in the reference (it has no gradients). */
int lp2d_propagate_adjoint_ws(int n, int batch, int dtype, const void* grad_out,
                              const void* field_in, const float* phase, int64_t phase_stride_b,
                              const lp2d_optics* optics, void* grad_in, float* grad_phase,
                              int adj_flags, void* work, void* stream);

/* lp2d_propagate_adjoint for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference (it has no gradients); same arguments and results as
 * lp2d_propagate_adjoint (no workspace: the intermediate plane stays in grad_in),
 * except that with a shared screen (phase_stride_b = 0) over a batch, grad_phase is a
 * single n x n plane receiving sum_b gphi[b]; the per-field or per-group planes of
 * that sum live in a 16 MiB device array of the library (one per device), so the call
 * allocates nothing.  Limits: those planes must fit in 16 MiB (batch x n^2 floats for
 * sizes that run in one whole-field launch, e.g. batch <= 256 at 128^2 complex64;
 * otherwise about 8 elements x 512K threads: every size fits), else LP2D_E_SIZE with
 * the figure.  lp2d_init(device) must have run (LP2D_E_INIT otherwise; it is what
 * allocates).  Calls on one device share that array: two of them that overlap in time
 * (two streams, two threads) race and must be ordered by the caller.  The array keeps
 * the last call's planes; every plane a call reads it has written first. */
int lp2d_propagate_adjoint_noalloc(int n, int batch, int dtype, const void* grad_out,
                                   const void* field_in, const float* phase,
                                   int64_t phase_stride_b, const lp2d_optics* optics,
                                   void* grad_in, float* grad_phase, int adj_flags,
                                   void* stream);

/* Phase-mask gradient as a separate pass, for callers that compose their own adjoint
 * from lp2d_fft2 / lp2d_spectral_apply:
 *   gphi[b] = Im(conj(field[b] * exp(i*phase)) * grad[b])
 * (the epilogue lp2d_propagate_adjoint fuses into its last sweep; grad is g1 there).
 * phase may be NULL (no factor); phase_stride_b is 0 (one screen shared by the batch)
 * or n*n.  sum_batch != 0: grad_phase is one (n, n) float32 plane receiving
 * sum_b gphi[b]; otherwise (batch, n, n).  No counterpart in the reference (it has no
 * gradients). */
int lp2d_phase_grad(int n, int batch, int dtype, const void* field, const void* grad,
                    const float* phase, int64_t phase_stride_b, float* grad_phase,
                    int sum_batch, void* stream);

/* Sum `batch` float planes of `count` elements: out[i] = sum_b in[b*count + i].
 * Used to reduce per-field phase gradients of a shared mask. */
int lp2d_sum_planes(int64_t count, int batch, const float* in, float* out, void* stream);

/* ---------------------------------------------------------------- plans */
/* A plan fixes one field geometry (n, batch, dtype) on the device current at
 * lp2d_plan_create, initialises that device (the allocation point: the twiddle
 * tables), and keeps a copy of the kernel schedule then selected for n
 * (lp2d_set_variant / lp2d_set_chunk, or the autotuner's table): later selections
 * change nothing a plan runs, so plans used from several threads each run what they
 * were made with.  A plan call must run with the plan's device current (LP2D_E_ARG
 * otherwise), takes the same arguments as the plan-less entry point of the same name
 * minus (n, batch, dtype), and allocates nothing; the adjoint forms take their
 * workspace from the caller (lp2d_plan_workspace_size).  No counterpart in the
 * reference, whose backend is selected per Beam2D (beam2d.py:105-109). */
typedef struct lp2d_plan lp2d_plan;
#define LP2D_OP_FFT2 0
#define LP2D_OP_PROPAGATE 1
#define LP2D_OP_ADJOINT 2
#define LP2D_OP_STACK 3
#define LP2D_OP_STACK_ADJOINT 4
int lp2d_plan_create(int n, int batch, int dtype, lp2d_plan** out);
int lp2d_plan_destroy(lp2d_plan* plan);
/* Bytes of `work` operation `op` (LP2D_OP_*) needs on this plan (0: none; -1: bad
 * plan or op). */
int64_t lp2d_plan_workspace_size(const lp2d_plan* plan, int op);
/* lp2d_fft2 (beam2d.py:207-229). */
int lp2d_plan_fft2(const lp2d_plan* plan, const void* in, void* out, int inverse, void* stream);
/* lp2d_propagate (beam2d.py:339-342, 481-490, 370-373). */
int lp2d_plan_propagate(const lp2d_plan* plan, const void* field_in, void* field_out,
                        const float* phase, int64_t phase_stride_b, const lp2d_optics* optics,
                        void* spectrum_out, void* stream);
/* lp2d_propagate_adjoint with the caller's workspace: `work` holds at least
 * lp2d_plan_workspace_size(plan, LP2D_OP_ADJOINT) bytes (LP2D_E_SIZE otherwise; NULL
 * when that is 0), must not alias the other arrays, and is undefined on return.  It
 * carries the column pass's output (complex64 three-sweep sizes) and, with a shared
 * screen (phase_stride_b = 0) over a batch, the per-group gradient planes, so that
 * grad_phase is then ONE n x n plane receiving sum_b gphi[b] (per-field phase:
 * (batch, n, n) as in lp2d_propagate_adjoint). */
int lp2d_plan_propagate_adjoint(const lp2d_plan* plan, const void* grad_out,
                                const void* field_in, const float* phase,
                                int64_t phase_stride_b, const lp2d_optics* optics,
                                void* grad_in, float* grad_phase, int adj_flags, void* work,
                                int64_t work_bytes, void* stream);
/* lp2d_propagate_stack (beam2d.py:339-342, 481-490 per plane). */
int lp2d_plan_propagate_stack(const lp2d_plan* plan, const void* field_in, void* field_out,
                              const float* screens, int nscreens, const lp2d_optics* optics,
                              void* const* planes_out, void* stream);
/* lp2d_propagate_stack_adjoint; work of lp2d_plan_workspace_size(plan,
 * LP2D_OP_STACK_ADJOINT) bytes. */
int lp2d_plan_propagate_stack_adjoint(const lp2d_plan* plan, const void* grad_out,
                                      const void* const* planes_in, const float* screens,
                                      int nscreens, const lp2d_optics* optics, void* grad_in,
                                      float* grad_screens, int adj_flags, void* work,
                                      int64_t work_bytes, void* stream);

/* ---------------------------------------------------------------- uneven batches */
/* One independent field of an uneven batch: its own size, distance and optics. */
typedef struct lp2d_field_desc {
  int32_t n;
  int32_t reserved;
  const void* src;        /* (n, n) field in  */
  void* dst;              /* (n, n) field out (may equal src) */
  double area_size, wl, z, aperture_radius;
} lp2d_field_desc;

typedef struct lp2d_batch lp2d_batch;

/* Groups the fields by size and uploads, per size, device tables of their source and
 * destination pointers and optics (allocates; call outside graph capture).  dtype
 * applies to all fields; 2 <= n <= 8192, powers of two. */
int lp2d_batch_create(const lp2d_field_desc* descs, int count, int dtype, lp2d_batch** out);
/* Propagates every field src -> dst: per size, the tuned kernels of that size read
 * the tables (one launch when the whole field fits in LDS, three sweeps otherwise),
 * largest size first.  Allocates nothing, so it can be captured in a hipGraph.  Must
 * run on the device the batch was created on. */
int lp2d_batch_run(lp2d_batch* batch, void* stream);
int lp2d_batch_destroy(lp2d_batch* batch);

/* ---------------------------------------------------------------- tracing */
/* Kernel timing on the launch stream: while enabled, every kernel launch of the
 * library is bracketed by a pair of HIP events recorded on the stream it runs on.
 * lp2d_trace_collect waits for them, sums the elapsed time per kernel name and
 * clears the record.  Used by bench.py for per-kernel roofline figures.  No
 * counterpart in the reference. */
typedef struct lp2d_trace_stat {
  char name[64];
  int32_t launches;
  int32_t reserved;
  double total_ms;
} lp2d_trace_stat;
int lp2d_trace_enable(int on);
int lp2d_trace_collect(lp2d_trace_stat* stats, int max_stats, int* n_stats);

/* ---------------------------------------------------------------- tuning */
/* Number of kernel variants (schedules) available for (n, dtype, kind), kind 0 =
 * row pass, 1 = column pass, 2 = whole-field-in-LDS kernel (0 or 1: sizes that use
 * it run every operation in one launch), 3 = chained-plane row pass of
 * lp2d_propagate_stack (0 where the plane is never block-tiled); and the one in use
 * (kinds 0, 1, 3).  Used by the autotuner. */
int lp2d_variant_count(int n, int dtype, int kind);
int lp2d_set_variant(int n, int dtype, int kind, int variant);
int lp2d_get_variant(int n, int dtype, int kind);
/* Fields per chunk of the three-sweep field-to-field propagation at size n (0, the
 * default, runs the whole batch per sweep).  A chunk's intermediate planes stay in
 * the Infinity Cache between its sweeps.  Used by the autotuner. */
int lp2d_set_chunk(int n, int dtype, int fields);
int lp2d_get_chunk(int n, int dtype);
/* Describes a variant as "E=16 W=16 ..." into buf (returns length). */
int lp2d_variant_name(int n, int dtype, int kind, int variant, char* buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif /* LP2D_H */
