/*
 * gpe_hip.h -- C ABI of libgpe_hip.so, the MI355X (gfx950) implementation of
 * the Strang-splitting time step behind
 *     solve(prob, StrangSplitting(), tspan; dt, nsaves)
 * of GeneralizedGrossPitaevskii.jl.
 *
 * Conventions (identical to Julia's memory layout, so a Julia caller passes
 * its arrays unchanged):
 *   - every field is a column-major array, dimension 0 contiguous, complex
 *     values interleaved (re, im);
 *   - the M components are separate arrays (the reference's NTuple{M} u0);
 *   - dimensions beyond the `ndim` spatial ones are batch (trajectory)
 *     dimensions, flattened into `batch`;
 *   - host-side tables (dispersion, potential, pump) are complex FP64;
 *   - pointers to fields and results may be host or device memory.
 * Every function returns GPE_OK (0) or a negative GPE_ERR_* code; the text of
 * the last error of the calling thread is returned by gpe_last_error().
 * A handle is not re-entrant; distinct handles may be used concurrently.
 */
#ifndef GPE_HIP_H
#define GPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPE_ABI_VERSION 1

enum gpe_status {
  GPE_OK = 0,
  GPE_ERR_INVALID = -1,     /* bad argument (shape, size, pointer, value)      */
  GPE_ERR_UNSUPPORTED = -2, /* valid in the reference, not supported here      */
  GPE_ERR_DEVICE = -3,      /* HIP runtime error                               */
  GPE_ERR_STATE = -4,       /* call out of order (e.g. envelope length)        */
  GPE_ERR_NOMEM = -5        /* device allocation failed                        */
};

enum gpe_dtype { GPE_C64 = 0 /* ComplexF32 fields */, GPE_C128 = 1 /* ComplexF64 fields */ };

/* Static-array form of a per-point term, as returned by the reference's user
 * closures: a scalar, an SVector{M} (elementwise) or an SMatrix{M,M}. */
enum gpe_form { GPE_FORM_NONE = 0, GPE_FORM_SCALAR = 1, GPE_FORM_DIAG = 2, GPE_FORM_FULL = 3 };

enum gpe_pump_kind { GPE_PUMP_NONE = 0, GPE_PUMP_TABLE = 1, GPE_PUMP_SEPARABLE = 2, GPE_PUMP_SAMPLES = 3 };

enum gpe_noise_kind { GPE_NOISE_NONE = 0, GPE_NOISE_COMPLEX = 1, GPE_NOISE_REAL = 2 };

typedef struct gpe_problem_desc {
  int32_t ndim;   /* spatial dimensions: 1, 2 or 3 (length(lengths))          */
  int32_t ncomp;  /* M components: 1..4 (length(u0)); M*n[d]/8 <= 1024, LDS   */
  int64_t n[3];   /* grid points per spatial dimension (powers of two, 16..4096) */
  double L[3];    /* box lengths (prob.lengths)                               */
  int64_t batch;  /* product of the trailing (trajectory) dimensions, >= 1    */
  int32_t dtype;  /* enum gpe_dtype                                           */
  int32_t device; /* HIP device ordinal                                       */
} gpe_problem_desc;

typedef struct gpe_handle gpe_handle;

/* Version of this ABI (GPE_ABI_VERSION). */
int gpe_version(void);

/* Text of the last error on the calling thread ("" if none). */
const char* gpe_last_error(void);

/* Allocate device state for one problem shape.  Replaces the allocation half
 * of init(prob, ::StrangSplitting, tspan) (src/strang_splitting.jl:41-49) and
 * the checks of GrossPitaevskiiProblem (src/problem.jl:112-115). */
int gpe_create(const gpe_problem_desc* desc, gpe_handle** out);

/* Release everything owned by the handle. */
int gpe_destroy(gpe_handle* h);

/* Dispersion D(k) sampled on the reciprocal grid (src/problem.jl:135-139) in
 * FFT order, column-major over the spatial dims, complex FP64:
 *   GPE_FORM_SCALAR: 1 value per k point
 *   GPE_FORM_DIAG:   M values per k point (component fastest)
 *   GPE_FORM_FULL:   M*M values per k point (column-major matrix); for M > 2
 *                    a matrix whose off-diagonal entries are all zero is applied
 *                    as its diagonal, any other is exponentiated per k in FP64
 *                    (scaling and squaring) and couples all M components
 *   GPE_FORM_NONE:   no dispersion (D = AdditiveIdentity), D must be NULL.
 * gpe_solve turns it into exp(-i dt D(k)) on the device (per-k matrix
 * exponential for FULL).  Replaces get_exponential(dispersion, ...)
 * (src/misc.jl:12-20, src/strang_splitting.jl:53). */
int gpe_set_dispersion(gpe_handle* h, int32_t form, const double* D);

/* Potential V(r) on the direct grid (src/problem.jl:129-133), SCALAR or DIAG
 * form as above; exp(-i dt/2 V) is formed on the device.  Replaces
 * get_exponential(potential, ...) (src/strang_splitting.jl:54). */
int gpe_set_potential(gpe_handle* h, int32_t form, const double* V);

/* Nonlinearity G(u) of the descriptor family
 *     G_i(u) = c[i] + sum_j g[i + M*j] |u_j|^2        (complex FP64)
 * which the real-space half-step applies as cis(-dt/2 G(u))
 * (src/kernels.jl:44).  g == NULL and c == NULL clears it. */
int gpe_set_nonlinearity(gpe_handle* h, const double* g, const double* c);

/* Matrix nonlinearity (two components): nonlinearity(u, param) returning a
 * 2x2 SMatrix, which the real-space half-step applies as the matrix
 * exponential cis(-dt/2 G(u)) to the component vector (src/kernels.jl:23,44;
 * LinearAlgebra's cis of a matrix), in the descriptor form
 *     G_ab(u) = c[a + 2b] + sum_j g[(a + 2b) + 4j] |u_j|^2     (complex FP64).
 * Replaces the vector form until gpe_set_nonlinearity is called again.
 * With a potential the reference's product rules apply (src/kernels.jl:9-15):
 * a scalar V scales the matrix, an SVector V gives the vector
 * exp(-i dt/2 G) exp(-i dt/2 V), applied elementwise.  Noise is drawn as for
 * the vector form.  The half-steps run in a per-point kernel between
 * transform-only passes (2D, 3D) or k-space-only whole steps (1D). */
int gpe_set_nonlinearity_matrix(gpe_handle* h, const double* g, const double* c);

/* Pump F(r,t) = profile(r) * envelope(t) (src/misc.jl:22-42).
 *   GPE_PUMP_TABLE:     profile = M blocks of prod(n) complex values.
 *   GPE_PUMP_SEPARABLE (1D and 2D): profile = M blocks of (n[0] + ... + n[ndim-1])
 *                       complex values; F_c(r) = prod_d p_{c,d}[i_d].
 * envelope: n_env complex FP64 samples at the times gpe_pump_times() lists
 * for the next gpe_solve; n_env == 0 means envelope == 1 (static pump).
 *   GPE_PUMP_SAMPLES:   F(r,t) that does not factor, sampled on the grid at
 *                       every pump time (what evaluate_pump! of
 *                       src/misc.jl:34-42 computes before each half-step):
 *                       profile = M blocks of n_env planes of prod(n) complex
 *                       values, plane j at the j-th time gpe_pump_times()
 *                       lists; envelope must be NULL.  The planes live in
 *                       device memory (at most 32 GiB per component) and cover
 *                       one gpe_solve: gpe_advance and gpe_steady_state
 *                       reject them, and the 2D step loop is not replayed
 *                       as a graph. */
int gpe_set_pump(gpe_handle* h, int32_t kind, const double* profile, const double* envelope,
                 int64_t n_env);

/* Position noise -i sqrt(dt/2) eta_c xi_c per half-step with constant eta
 * (position_noise_func, src/kernels.jl:42) and xi drawn in-kernel from
 * Philox4x32-10 with key `seed`; `call0` is the counter of the next draw (two
 * draws per step).  kind GPE_NOISE_NONE disables noise. */
int gpe_set_noise(gpe_handle* h, int32_t kind, const double* eta, uint64_t seed, uint64_t call0);

/* Ensemble observables of every saved frame of the next gpe_solve calls,
 * reduced on the device over the batch (trajectory) dimension in FP64 with
 * compensated summation: for component c, frame f and spatial point r
 *     n [(c*nframes + f)*prod(n) + r] = mean_b |u_c(r, b)|^2
 *     n2[(c*nframes + f)*prod(n) + r] = mean_b |u_c(r, b)|^4
 * (host or device memory, nframes = nsaves + save_start).  The reference has
 * no such entry point: its users compute these moments from the returned
 * frames (examples/truncated_wigner.jl:110-165, test/windowed_ft.jl:31-49).
 * n == NULL disables. */
int gpe_set_observables(gpe_handle* h, double* n, double* n2);

/* resolve_fixed_timestepping (src/fixed_time_stepping.jl:14-24). */
int gpe_resolve_timestepping(double dt, double t0, double t1, int64_t nsaves, double* dt_out,
                             int64_t* steps_per_save);

/* The times at which solve evaluates the pump (src/strang_splitting.jl:58,
 * 81, 87, 89 with t accumulated as in src/fixed_time_stepping.jl:44):
 * times[0] = t0, and for step s = 1..S (t_s = t0 + dt + ... + dt, s terms)
 * times[2s-1] = t_s + dt/2, times[2s] = t_s + dt.  n must be 1 + 2*S. */
int gpe_pump_times(double t0, double t1, double dt, int64_t nsaves, double* times, int64_t n);

/* The whole fixed-step solve: init + solve! (src/strang_splitting.jl:32-67,
 * src/fixed_time_stepping.jl:26-54, src/strang_splitting.jl:86-90).
 *   u0[c]     : M pointers, field arrays (prod(n) * batch complex values)
 *   result[c] : M pointers, arrays of shape (n..., batch, nsaves + save_start);
 *               result == NULL skips the frames (observables-only runs)
 *   ts        : nsaves + save_start doubles (may be NULL)
 * Returns after result and ts are written. */
int gpe_solve(gpe_handle* h, const void* const* u0, double t0, double t1, double dt,
              int64_t nsaves, int32_t save_start, void* const* result, double* ts);

/* How gpe_solve issues the 2D step loop.  The reference steps on the host,
 * one broadcast kernel per operation (src/fixed_time_stepping.jl:40-46,
 * src/strang_splitting.jl:86-90); the choice has no counterpart there.
 *   GPE_SCHEDULE_LAUNCHES: two launches per step (column pass, row pass).
 *   GPE_SCHEDULE_GRAPH:    the steps inside each save block are built once
 *                          per solve as a HIP graph and replayed per block.
 *                          Needs no noise, a static envelope and >= 2 steps
 *                          per save (GPE_ERR_UNSUPPORTED otherwise).
 *   GPE_SCHEDULE_AUTO:     the graph when eligible and nsaves >= 8 (default).
 * gpe_steady_state replays the steps between residual checks the same way
 * (no noise, check_every >= 4) unless the schedule is LAUNCHES.  The kernels
 * and their arguments are the same under every schedule, so the results are
 * identical bit for bit. */
enum gpe_schedule { GPE_SCHEDULE_AUTO = 0, GPE_SCHEDULE_LAUNCHES = 1, GPE_SCHEDULE_GRAPH = 2 };
int gpe_set_schedule(gpe_handle* h, int32_t schedule);

/* The schedule the last 2D gpe_solve or gpe_steady_state ran with
 * (GPE_SCHEDULE_LAUNCHES or GPE_SCHEDULE_GRAPH). */
int gpe_get_schedule(gpe_handle* h, int32_t* schedule);

/* Planner of the 2D passes' launch shapes, the counterpart of the FFTW
 * planner behind get_fft_plans (src/misc.jl:53-58, FFTW.jl's default flags):
 *   GPE_PLAN_ESTIMATE: the built-in shapes.
 *   GPE_PLAN_MEASURE:  the next gpe_solve / gpe_steady_state of a 2D handle
 *                      first times whole steps for every valid pair of
 *                      row-pass shape (rows per workgroup) and column-pass
 *                      shape (columns per workgroup) on this handle, with its
 *                      tables and terms, and keeps the fastest pair; it
 *                      measures again when the terms that select kernel code
 *                      (noise, nonlinearity, dispersion, potential or pump
 *                      kind) change.  Measured plans are kept for the process
 *                      (per device, shape, precision, batch and terms).
 *   GPE_PLAN_AUTO:     (default) a plan the process measured before for the
 *                      same key; else MEASURE when the call runs >= 1000
 *                      steps (the measurement takes ~25 ms); else ESTIMATE.
 * Setting the planner drops any plan fixed by gpe_set_plan.  No shape changes
 * the arithmetic: results are identical bit for bit under every plan. */
enum gpe_planner { GPE_PLAN_ESTIMATE = 0, GPE_PLAN_MEASURE = 1, GPE_PLAN_AUTO = 2 };
int gpe_set_planner(gpe_handle* h, int32_t planner);

/* Fix the launch shapes (e.g. a plan measured before, the counterpart of
 * FFTW wisdom): x_rows rows per workgroup of the row pass and y_cols columns
 * per workgroup of the column pass, 0 for the built-in shape.  2D only.
 * GPE_ERR_INVALID for a shape the kernels do not provide. */
int gpe_set_plan(gpe_handle* h, int32_t x_rows, int32_t y_cols);

/* The shapes in use.  cand[12] and us[6] (both may be NULL) receive the
 * (x_rows, y_cols) pairs of the last measurement and their mean time per
 * step in us (a column pass and a row pass); unused entries are 0. */
int gpe_get_plan(gpe_handle* h, int32_t* x_rows, int32_t* y_cols, int32_t* cand, double* us);

/* Momentum-space occupation and pair correlation of a 1D truncated-Wigner
 * ensemble frame, the observables of examples/truncated_wigner.jl:110-165:
 * with ft_b = fft(u_b)/N per trajectory b (one component of one saved frame,
 * `frame` = N x batch values in the field precision, host or device),
 *   nk[k]        = mean_b |ft_b(k)|^2 - commutator/2
 *   g2[k' N + k] = mean_b |ft_b(k)|^2 |ft_b(k')|^2
 *                  - (1 + delta_kk') commutator/2 (nbar(k) + nbar(k') - commutator/2)
 * (the example's G2k; g2(k,k') = G2 / (n(k) n(k')) is left to the caller),
 * nbar = mean_b |ft_b|^2, all in FFT order (the example's fftshift is a
 * permutation), FP64 with compensated trajectory sums.  commutator is 1/L
 * in the example.  nk (N doubles) and g2 (N*N doubles) may be host or device
 * memory.  1D handles only. */
int gpe_kspace_g2(gpe_handle* h, const void* frame, double commutator, double* nk, double* g2);

/* Windowed first-order correlation of a 1D ensemble frame, the `correlation`
 * of test/windowed_ft.jl:31-49: with X_j,b = fft(w_j .* u_b) (unnormalised,
 * FFT order) for the windows w1, w2 (N complex FP64 values each, host or
 * device) and the trajectories u_b of `frame` (N x batch values, field
 * precision, host or device),
 *   g1[2 (k2 N + k1) + {0, 1}] = sum_b X_1,b(k1) conj(X_2,b(k2)) / (N batch)
 * (complex FP64, host or device; compensated sums).  The reference's
 * fftshift / ifftshift pair maps this to its layout by the index shift
 * k -> (k + N/2) mod N and the sign (-1)^k per index.  1D handles only. */
int gpe_windowed_g1(gpe_handle* h, const void* frame, const double* w1, const double* w2, double* g1);

/* Normal-ordered density-density correlation along dimension 0 of an
 * ensemble frame (one component, `frame` = prod(n) x batch values in the
 * field precision, host or device), averaged over x, the other dimensions and
 * the trajectories, with the ordering correction of
 * examples/truncated_wigner.jl:140-157 (delta = [r == 0]):
 *   g2r[r] = < |u(x)|^2 |u(x + r)|^2 > - (1 + delta) c/2 (2 nw - c/2),
 *   *nphys = nw - c/2,   nw = < |u(x)|^2 >,   r = 0 .. n[0]-1 (periodic),
 * c = the commutator per grid point (1/dA, 1/dx).  g2(r) of a homogeneous
 * state is g2r[r] / nphys^2.  FP64 with compensated sums; the
 * autocorrelation of each line is formed through the FFT.  Used for the
 * truncated-Wigner ensembles' g2(r) without moving frames to the host. */
int gpe_density_g2_x(gpe_handle* h, const void* frame, double commutator, double* g2r, double* nphys);

/* Copy the field after the last solve (real space) into u[c]. */
int gpe_get_field(gpe_handle* h, void* const* u);

/* Device time in ms of the step loop of the last gpe_solve / gpe_advance,
 * from HIP events recorded on the handle's stream around the loop. */
int gpe_last_solve_ms(gpe_handle* h, float* ms);

/* Continue the last solve's configuration `nsteps` more steps from the
 * current field (envelope frozen at its last sample), no frame output.
 * *ms (may be NULL) receives the device time of the loop. */
int gpe_advance(gpe_handle* h, int64_t nsteps, float* ms);

/* As gpe_advance, with HIP events around every kernel launch: *ms_x is the
 * mean event-bracketed time in ms of the fused real-space kernel (2D x-pass,
 * or the 1D whole-step kernel), *ms_y that of the 2D k-space y-pass (0 in
 * 1D), and *ms_event (may be NULL) the mean time of an empty event pair on
 * the same stream, the bracketing's own latency. */
int gpe_profile_kernels(gpe_handle* h, int64_t nsteps, double* ms_x, double* ms_y, double* ms_event);

/* Steady-state search.  The reference has no convergence check (users solve
 * over a long tspan, src/fixed_time_stepping.jl:26-54); this steps from u0
 * with the given dt (static pump envelope required) and every check_every
 * steps (>= 2) forms r = ||u_{n+1} - u_n||_2 / ||u_{n+1}||_2 over all points
 * and components in FP64 with compensated sums, from two consecutive
 * real-space fields.  It stops at r <= tol or when max_steps would be
 * exceeded.  *steps and *residual receive the steps taken and the last r;
 * history (may be NULL) receives r of every check (at most max_hist, count in
 * *nhist); *ms (may be NULL) the device time.  The final real-space field
 * stays in the handle (gpe_get_field). */
int gpe_steady_state(gpe_handle* h, const void* const* u0, double dt, double tol, int64_t check_every,
                     int64_t max_steps, int64_t* steps, double* residual, double* history, int64_t max_hist,
                     int64_t* nhist, float* ms);

/* ----------------------------------------------------------------------
 * Uneven-size instance sweeps.  A sweep is a set of independent 1D,
 * single-component problems of different grid sizes n[i] (powers of two
 * 16..4096) and lengths L[i], each with its own D(k), G and pump profile and
 * a shared pump envelope, advanced together on one GPU.  It replaces the
 * reference's way of running such a sweep, one solve per
 * GrossPitaevskiiProblem (src/fixed_time_stepping.jl:79-81) in a loop.
 * Concatenated layout: instance i occupies points off[i] .. off[i]+n[i]-1,
 * off[i] = n[0] + ... + n[i-1].
 * ---------------------------------------------------------------------- */
typedef struct gpe_sweep gpe_sweep;

int gpe_sweep_create(int64_t ninst, const int64_t* n, const double* L, int32_t dtype, int32_t device,
                     gpe_sweep** out);
int gpe_sweep_destroy(gpe_sweep* s);

/* Concatenated D_i(k) on each instance's reciprocal grid (FFT order),
 * scalar form, complex FP64 (sum n[i] values). */
int gpe_sweep_set_dispersion(gpe_sweep* s, const double* D);

/* Per-instance G_i(u) = c[i] + g[i] |u|^2 (complex FP64, ninst values each). */
int gpe_sweep_set_nonlinearity(gpe_sweep* s, const double* g, const double* c);

/* Concatenated pump profiles (complex FP64, sum n[i] values; an all-zero
 * block leaves that instance unpumped) times one envelope shared by all
 * instances, sampled at the times gpe_pump_times lists (n_env == 0: 1). */
int gpe_sweep_set_pump(gpe_sweep* s, const double* profile, const double* envelope, int64_t n_env);

/* Per-instance constant noise amplitude eta[i] with Philox key `seed`. */
int gpe_sweep_set_noise(gpe_sweep* s, int32_t kind, const double* eta, uint64_t seed, uint64_t call0);

/* The fixed-step solve of every instance.  u0: concatenated initial fields.
 * result: instance-major frames, instance i's block of n[i] * nframes values
 * (one column-major (n[i], nframes) array, the reference's result shape) at
 * offset off[i] * nframes, nframes = nsaves + save_start; NULL skips frames.
 * u0/result may be host or device memory; ts as in gpe_solve. */
int gpe_sweep_solve(gpe_sweep* s, const void* u0, double t0, double t1, double dt, int64_t nsaves,
                    int32_t save_start, void* result, double* ts);

/* Device time in ms of the step loop of the last gpe_sweep_solve. */
int gpe_sweep_last_ms(gpe_sweep* s, float* ms);

/* ----------------------------------------------------------------------
 * Kernels of the Fourier-neural-operator warm start for steady states.  The
 * reference has no learned component; these serve the surrogate whose output
 * seeds gpe_steady_state.  All arrays are device FP32, row-major; `stream` is
 * a hipStream_t (NULL: the default stream).
 * ---------------------------------------------------------------------- */

/* C[b] = alpha A[b] B[b] + beta C[b] for b < batch (<= 65535): A is m x k,
 * B k x n, C m x n, element (i, j) of each at base + i*rs + j*cs and batch b
 * at + b*bs, with strides = {A rs, A cs, B rs, B cs, C rs, C cs, A bs, B bs,
 * C bs} in elements (transposes are stride swaps).  Operands are rounded to
 * BF16 and multiplied on the matrix cores with FP32 accumulation. */
int gpe_fno_gemm(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k,
                 const int64_t* strides, int64_t batch, float alpha, float beta, void* stream);

/* gpe_fno_gemm with FP32 operands (FP32 matrix-core products): the
 * transforms that band-limit a prediction, where BF16-rounded basis functions
 * would add error at every wavenumber.  m / 64 <= 65535 for both. */
int gpe_fno_gemm_f32(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k,
                     const int64_t* strides, int64_t batch, float alpha, float beta, void* stream);

/* y = act(x + bias[ch]) over rows x cols values, ch = column (channels_last)
 * or (row mod C); act 1 = GELU (tanh form), 0 = identity; pre (may be NULL)
 * keeps x + bias for the backward. */
int gpe_fno_bias_act(const float* x, const float* bias, float* pre, float* y, int64_t rows, int64_t cols,
                     int64_t C, int32_t channels_last, int32_t act, void* stream);

/* dx = dy * GELU'(pre). */
int gpe_fno_act_backward(const float* pre, const float* dy, float* dx, int64_t n, void* stream);

/* y = a x + b y over n values. */
int gpe_fno_axpby(const float* x, float* y, float a, float b, int64_t n, void* stream);

/* One Adam step (bias-corrected with `step` >= 1) on n FP32 parameters. */
int gpe_fno_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                 float eps, int64_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GPE_HIP_H */
