/*
 * macstokes.h -- C ABI of the MI355X (gfx950) MAC shifted-Stokes stage solve.
 *
 * The library solves, for each right-hand-side column f,
 *     (mass*I + viscosity*K) u + G p = f,   D u = 0,   mean(p) = 0
 * on a uniform no-slip MAC grid (the reference's `TensorStokes.solve`,
 * solver/mac/tensor_stokes.py:245-364): the velocity Helmholtz blocks are inverted
 * exactly by 2-D DST transforms written as MFMA GEMMs, and the pressure Schur
 * complement D L^-1 D^T is solved by preconditioned CG with the preconditioner
 * mass*(D D^T)^+ + viscosity*I (2-D DCT-II), all on the GPU.
 *
 * Layouts (the reference's, solver/mac/grid.py:52-65 and kernels.py:55):
 *   velocity column: size = (nx-1)*ny + nx*(ny-1) doubles,
 *                    [u interior faces (ny, nx-1) row-major ; v interior faces (ny-1, nx)]
 *   pressure column: np = nx*ny doubles, cell (j, i) at j*nx + i
 *   k columns: column c starts at base + c*ld (ld >= size or >= np).
 *
 * Every call is synchronous: it returns after the GPU work is done, and keeps no
 * pointer it was given.  Pointers are host memory unless `device_pointers` is 1,
 * then they are device memory of the handle's device.  One handle must not be used
 * from two threads at once.
 *
 * Return codes (mapped by the Python binding to the reference's exceptions,
 * tensor_stokes.py:137-147, 246-257, 322-323, 345-351):
 */
#ifndef MACSTOKES_H
#define MACSTOKES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAC_OK 0
#define MAC_INVALID_ARGUMENT 1 /* ValueError: invalid controls, shifts or RHS    */
#define MAC_NOT_CONVERGED 2    /* RuntimeError: Schur CG hit max_iterations       */
#define MAC_GATE_FAILED 3      /* RuntimeError: residual / divergence gate failed */
#define MAC_HIP_ERROR 4        /* RuntimeError: HIP runtime failure               */

#define MAC_PRECISION_FP64 0   /* FP64 MFMA transforms (the plain GPU solve)       */
#define MAC_PRECISION_MIXED 1  /* BF16 MFMA transforms, FP32 accumulation, inside
                                  an FP64 defect-correction loop                   */

typedef struct mac_stokes_opts {
  double tolerance;        /* residual / divergence gate (reference default 1e-9)    */
  double krylov_tolerance; /* Schur-CG rtol, ||r|| < rtol*||b|| (default 1e-12)      */
  int32_t max_iterations;  /* Schur-CG maxiter (default 100)                         */
  int32_t precision;       /* MAC_PRECISION_*                                        */
  double refine_tolerance; /* MIXED: stop when the residual and max|D u| are both
                              <= this (default 1e-10)                               */
  double inner_tolerance;  /* MIXED: cap of the inner Schur-CG rtol of a sweep, which
                              is 0.25 * refine_tolerance / (the sweep's defect) clamped
                              to [min(1e-4, cap), cap]; FP64 refinement: its sweeps'
                              rtol (default 1e-3)                                    */
  int32_t max_sweeps;      /* MIXED: outer defect-correction sweep cap (default 30)  */
  int32_t device;          /* HIP device ordinal                                      */
  int32_t mixed_split;     /* MIXED: bit 0 -> velocity inverses, bit 1 -> pressure
                              preconditioner use the split bf16x3 GEMM (hi/lo parts
                              of data and matrix, ~16-bit operands); bit 2 -> their
                              forward passes run as even/odd pairs of half the
                              contraction length; bit 3 -> where every transform
                              length allows it (FFT length a power of two, n <= 2048),
                              the split transforms run as FP32 FFT passes instead;
                              default 15                                            */
  int32_t fp64_refine;     /* FP64: after the reference solve, up to this many FP64
                              defect-correction sweeps (inner Schur-CG rtol
                              inner_tolerance) until the residual and max|D u| are
                              both <= refine_tolerance; 0 (default) = the reference
                              algorithm alone                                        */
} mac_stokes_opts;

typedef struct mac_solve_report {
  double residual;        /* max|mass u + visc K u + G p - f| / (1 + max|f|)      */
  double divergence_inf;  /* max|D u|                                              */
  double velocity_max;    /* max|u| (the divergence gate's scale)                  */
  double seconds;         /* wall time of the call, device-synchronised            */
  int32_t sweeps;         /* outer defect-correction sweeps (1 for FP64)           */
  int32_t total_iterations; /* Schur-CG iterations summed over columns and sweeps  */
  int32_t warm_started;   /* 1: a warm start (history / previous) supplied the start;
                              2: the network's correction entered it; 0: zero start  */
  int32_t pad0;
  double first_residual;   /* MIXED / refined FP64: residual and max|D u| of the first */
  double first_divergence; /* FP64 defect (the starting iterate's)                     */
} mac_solve_report;

/* Fills `opts` with the reference defaults (tensor_stokes.py:128-136). */
void mac_stokes_default_opts(mac_stokes_opts* opts);

/* Replaces `TensorStokes.__init__` (tensor_stokes.py:128-174): validates controls,
 * uploads the DST/DCT matrices and eigenvalues of an nx x ny grid on [0,lx]x[0,ly].
 * Returns NULL on failure; mac_stokes_create_error() then describes why. */
void* mac_stokes_create(int32_t nx, int32_t ny, double lx, double ly, const mac_stokes_opts* opts);
const char* mac_stokes_create_error(void);
void mac_stokes_destroy(void* handle);
/* Message of the last failing call on this handle ("" if none). */
const char* mac_stokes_last_error(void* handle);

/* Replaces `TensorStokes.solve(rhs, mass=, viscosity=)` (tensor_stokes.py:245-364).
 * rhs: ncols columns of `size`; velocity out: ncols columns of `size`; pressure
 * out: ncols columns of `np` (mean zero).  `pressure_guess` (nullable, ncols columns
 * of np) starts the Schur CG from a guess instead of zero.  iterations (nullable)
 * receives ncols Schur-CG counts (FP64: as scipy's cg callback counts them). */
int32_t mac_stokes_solve(void* handle, const double* rhs, int64_t ld_rhs, int32_t ncols, double mass,
                         double viscosity, double* velocity, int64_t ld_velocity, double* pressure,
                         int64_t ld_pressure, const double* pressure_guess, int64_t ld_guess,
                         int32_t* iterations, mac_solve_report* report, int32_t device_pointers);

/* No counterpart in the reference (its solve is a synchronous NumPy call): runs the
 * handle's work on the caller's HIP stream (`use` = 1; `stream` may be the null stream)
 * or on its own (`use` = 0, the default).  On the caller's stream a solve with device
 * pointers is ordered like the caller's other work: it returns once the host has the
 * solve's scalars, without waiting for the output copies queued after them, and needs no
 * synchronisation of that stream before the call.  Pending work on the previous stream
 * is finished first. */
int32_t mac_stokes_set_stream(void* handle, void* stream, int32_t use);

/* ---- measurement: HIP-event timing of the transform GEMM launches ---- */
typedef struct mac_kernel_profile {
  double gemm_ms;        /* summed device time of the transform-GEMM launches       */
  double gemm_flops;     /* their MFMA work, 2*M*N*K per issued product (three per
                            k-step in the split bf16x3 form, one otherwise)         */
  double gemm_bytes;     /* their algorithmic bytes: A, Bt and C elements once each  */
  int64_t gemm_launches; /* number of launches timed                                */
  double gemm_transform_flops; /* the transform work of the passes they compute:
                            2*M*N*n per pass over a length-n transform (one product;
                            an even/odd pair counts as the one pass it computes)    */
  double fft_ms;         /* the same for the FFT transform passes (mixed_split bit 3): */
  double fft_bytes;      /* device time, algorithmic bytes (each element read once and */
  int64_t fft_launches;  /* written once), launches, and the transform work of the    */
  double fft_transform_flops; /* passes, counted as for the GEMMs (2*rows*n*n per
                            length-n transform of `rows` rows)                      */
} mac_kernel_profile;

/* No counterpart in the reference (its counters are wall-clock only,
 * tensor_stokes.py:169-174): switches per-launch event timing on/off and resets it. */
void mac_stokes_set_profiling(void* handle, int32_t on);
int32_t mac_stokes_get_profile(void* handle, mac_kernel_profile* out);

/* ---- learned warm start (no counterpart in the reference; CG stays the judge) ----
 * The network maps [pool4(b), pool4(p_prev)] (normalised) to pool4(p - p_prev): b is the
 * Schur right-hand side, p_prev the previous solve's pressure of the same column; the
 * guess is p_prev + the bilinearly upsampled, rescaled output.
 * Served grids: nx, ny multiples of 32. */
void* mac_cnn_create(int32_t width, int32_t device);
const char* mac_cnn_create_error(void);
void mac_cnn_destroy(void* cnn);
const char* mac_cnn_last_error(void* cnn);
int64_t mac_cnn_num_params(void* cnn);
/* Parameters as FP32, per conv layer c0..c6: weights (co, ci, 3, 3) then bias (co). */
int32_t mac_cnn_set_params(void* cnn, const float* host_params);
int32_t mac_cnn_get_params(void* cnn, float* host_params);
/* Element counts (FP32 units) of the 29 activation / work buffers one forward and
 * backward pass of n inputs of H x W needs, for a network of `width` channels: the
 * plan the network sizes its buffers by before any pass (host arithmetic only, no
 * device call).  MAC_INVALID_ARGUMENT unless H, W are multiples of 8. */
int32_t mac_cnn_plan(int32_t width, int32_t n, int32_t H, int32_t W, int64_t* counts);
/* x: [n][2][H][W], y: [n][1][H][W] device FP32, H and W multiples of 8. */
int32_t mac_cnn_forward(void* cnn, const float* x, float* y, int32_t n, int32_t H, int32_t W);
/* Mean squared error against target [n][1][H][W] and its gradient (device, FP32). */
int32_t mac_cnn_loss_grad(void* cnn, const float* x, const float* target, int32_t n, int32_t H, int32_t W,
                          float* grad_out, float* loss_out);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on that loss. */
int32_t mac_cnn_train_step(void* cnn, const float* x, const float* target, int32_t n, int32_t H, int32_t W,
                           float lr, float* loss_out);
/* Starting pressure when the caller passes none:
 *   MAC_WARM_OFF       zero, as the reference;
 *   MAC_WARM_NETWORK   the attached network's correction of a base guess (the history
 *                      extrapolation when a period is set, else the previous solve's
 *                      pressures); alpha is the pressure / right-hand-side scale it was
 *                      trained with;
 *   MAC_WARM_PREVIOUS  the previous solve's pressures;
 *   MAC_WARM_HISTORY   the same stage of the last K steps extrapolated polynomially in
 *                      time (K = mac_stokes_set_history_steps, fewer while fewer are
 *                      recorded; 3 p[-P] - 3 p[-2P] + p[-3P] for K = 3 and equal steps), P
 *                      the stage solves per step (mac_stokes_set_warm_history); a MIXED
 *                      handle extrapolates the velocities the same way.
 * FP64 handles start the Schur CG from theta g, theta = (g, b)/(g, S g); MIXED handles
 * take g as the initial outer iterate of the defect correction. */
#define MAC_WARM_OFF 0
#define MAC_WARM_NETWORK 1
#define MAC_WARM_PREVIOUS 2
#define MAC_WARM_HISTORY 3
int32_t mac_stokes_set_warm_start(void* handle, int32_t mode, void* cnn, double alpha);
/* Forget the previous-stage pressures (a new run). */
int32_t mac_stokes_reset_history(void* handle);
/* Stage solves per step for MAC_WARM_HISTORY and the network's base (0..16; 0 = off);
 * forgets the recorded history.  sdirk3_mrsav: 4, sdirk2_mrsav: 2. */
int32_t mac_stokes_set_warm_history(void* handle, int32_t period);
/* Steps the history extrapolation reaches back, 1..6 (default 6: degree 5 in time);
 * forgets the recorded history. */
int32_t mac_stokes_set_history_steps(void* handle, int32_t steps);
/* MAC_WARM_NETWORK with a history period: the network corrects the extrapolation only
 * while fewer than `steps` steps are recorded (1..7, default 3); from then on the
 * extrapolation alone is the guess.  FP64 handles combine base g and correction c as
 * a g + beta c with the smallest starting Schur residual (never worse than g alone). */
int32_t mac_stokes_set_network_depth(void* handle, int32_t steps);
/* Trial bookkeeping for the history: rollback = 0 commits the solves recorded since the
 * last mark (an accepted step), rollback = 1 forgets them (a rejected or failed trial),
 * so guesses extrapolate accepted steps only. */
int32_t mac_stokes_history_mark(void* handle, int32_t rollback);
/* Schur right-hand sides b of the last FP64 solve, ncols columns of np (device). */
int32_t mac_stokes_last_schur_rhs(void* handle, double* out, int64_t ld, int32_t ncols);
/* Network inputs / targets / scales exactly as the solver prepares them (device):
 * x[c] = [pool4(b)/s_c, pool4(base)/r_c], t[c] = pool4(target)/(alpha s_c), with s_c, r_c
 * the rms of the pooled planes; `scale` receives s (ncols) then r (ncols).  `target` is
 * the quantity to learn, p - base. */
int32_t mac_warm_prepare(int32_t nx, int32_t ny, const double* b, int64_t ldb, const double* pprev, int64_t ldp,
                         const double* target, int64_t ldt, int32_t ncols, double alpha, float* x, float* t,
                         double* scale);

/* ---- single operators, device pointers only (parity tests and host callers) ---- */

/* D: packed velocity columns -> cell divergence (kernels.py:46-87). */
int32_t mac_apply_divergence(void* handle, const double* velocity, int64_t ld_velocity, int32_t ncols,
                             double* out, int64_t ld_out);
/* D^T: pressure columns -> packed faces (kernels.py:89-132). */
int32_t mac_apply_divergence_transpose(void* handle, const double* pressure, int64_t ld_pressure,
                                       int32_t ncols, double* out, int64_t ld_out);
/* (mass*I + viscosity*K)^-1 on packed velocity columns (tensor_stokes.py:200-215).
 * On a MAC_PRECISION_MIXED handle this and the next entry point use the bf16
 * transforms (FP32 accumulation), as the mixed solve's inner iterations do. */
int32_t mac_velocity_inverse(void* handle, const double* in, int64_t ld_in, int32_t ncols, double mass,
                             double viscosity, double* out, int64_t ld_out);
/* (D D^T)^+ on pressure columns (tensor_stokes.py:217-243). */
int32_t mac_pressure_poisson_inverse(void* handle, const double* in, int64_t ld_in, int32_t ncols,
                                     double* out, int64_t ld_out);

/* ---- device MAC model operations (the schemes' algebra between stage solves) ----
 * Device pointers; kernels run on the null stream.  mac_dot and mac_max_abs synchronise
 * and return a host scalar; the others return once enqueued.  Status codes as above;
 * mac_model_last_error() holds the calling thread's last message. */
const char* mac_model_last_error(void);
/* N(u) per packed velocity column (kernels.py:6-35, via mac_ns.py:22-25); bitwise the
 * reference loop. */
int32_t mac_convection(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, int64_t ld,
                       int32_t ncols, double* out, int64_t ld_out);
/* out = sum_k coefficients[k] * fields[k], accumulated left to right (mac_ns.py:40-45);
 * nterms <= 8, fields[] a host array of device pointers. */
int32_t mac_combine(int64_t n, int32_t nterms, const double* coefficients, const double* const* fields,
                    double* out);
/* K u, G p = -D^T p and D u as the reference's CSR products (operators.py:6-29). */
int32_t mac_apply_K(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, double* out);
int32_t mac_grid_gradient(int32_t nx, int32_t ny, double lx, double ly, const double* pressure, double* out);
int32_t mac_grid_divergence(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, double* out);
/* sum a*b (fixed-order two-level reduction) and max|a| (NaN reported as inf). */
int32_t mac_dot(int64_t n, const double* a, const double* b, double* result);
int32_t mac_max_abs(int64_t n, const double* a, double* result);
/* The same two reductions without waiting: mode 0 writes sum a*b, mode 1 max|a|, into
 * the device double *out, in the null stream's order (the reference computes these
 * inline, solver/mac/grid.py:67-68 and the schemes' max-norm gates,
 * solver/schemes/sdirk3_mrsav.py:35-57); callers read many such slots in one copy. */
int32_t mac_reduce_into(int64_t n, const double* a, const double* b, int32_t mode, double* out);

/* ---- allocation-free inner product ----
 * No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that forbid hipMalloc): mac_dot's
 * result (same kernels, same summation order, bitwise equal) with its partial sums in
 * a file-scope device array of 4160 doubles (33 KB per device, held for the life of the
 * process) instead of library-allocated scratch.  1 <= n <= 2^31 - 1, else
 * MAC_INVALID_ARGUMENT.  Synchronous (the result is a host double).  Calls that overlap
 * on one device share the array and race; mac_noalloc_last_error() holds the calling
 * thread's last message. */
#define MAC_NOALLOC_MAX_N 2147483647LL
int32_t mac_dot_noalloc(int64_t n, const double* a, const double* b, double* result);
const char* mac_noalloc_last_error(void);

/* ---- batched cavities ----
 * Replaces a host loop of `TensorStokes.solve` (tensor_stokes.py:245-364) over many
 * independent cavities, e.g. one `MACNavierStokes` per cavity stepped in lockstep: every
 * cavity keeps its own grid, (mass, viscosity) and columns, and its results are those of
 * the cavity solved alone.  Segments (cavity i, column c) are laid out cavity-major,
 * column-minor; mac_batch_layout gives each segment's first element in the velocity and
 * pressure buffers (plus the totals at index count*ncols).
 * precision MAC_PRECISION_FP64: every cavity's result is bitwise that of the cavity solved
 * alone.  MAC_PRECISION_MIXED: per cavity the mixed form's outer FP64 defect correction
 * (refine_tolerance, inner_tolerance, max_sweeps) around inner Schur-CG solves whose
 * transforms are split bf16 (hi/lo parts of data and matrix, FP32 accumulation) and whose
 * vectors, dot products, defects and updates are FP64; finished cavities drop out of the
 * sweeps on the device. */
void* mac_batch_create(int32_t count, const int32_t* nx, const int32_t* ny, const double* lx, const double* ly,
                       int32_t ncols, const mac_stokes_opts* opts);
const char* mac_batch_create_error(void);
void mac_batch_destroy(void* handle);
const char* mac_batch_last_error(void* handle);
int32_t mac_batch_layout(void* handle, int64_t* velocity_offsets, int64_t* pressure_offsets);
/* Launch shape.  use_graph 0: direct launches; 1: one captured CG iteration replayed as a
 * hipGraph; with either, iters_per_check (1..16) iterations are queued between host reads
 * of the convergence flags.  use_graph 2 (mixed form only): the stage graph, one captured
 * hipGraph launch per stage whose outer tests, inner set-ups, CG iterations and updates are
 * switched by device flags, iters_per_check CG iterations per round (4 rounds per launch;
 * a stage that needs more is launched again), one host read per launch.  The results do
 * not depend on the shape. */
int32_t mac_batch_set_launch(void* handle, int32_t use_graph, int32_t iters_per_check);
/* Outer sweeps of the last batched solve (mixed form; 1 for FP64), -1 for a null handle. */
int32_t mac_batch_last_sweeps(void* handle);
/* No counterpart in the reference (its solves start from zero).  Mixed form only: each
 * batched solve's outer iterate (u, p) starts from the Lagrange extrapolation in time of
 * the solves `period`, 2 period, ... back (the same stage of the last `steps` steps,
 * 1..6), fewer while fewer are recorded; period 0 switches it off and every call
 * restarts the record.  The step sizes come from the solves' viscosity / mass (cavity 0;
 * the cavities step together).  MAC_INVALID_ARGUMENT for an FP64 handle. */
int32_t mac_batch_set_history(void* handle, int32_t period, int32_t steps);
/* rollback 0 keeps the solves since the last mark in the history (an accepted step);
 * 1 forgets them (a rejected trial). */
int32_t mac_batch_history_mark(void* handle, int32_t rollback);
/* mass, viscosity: per cavity.  iterations: per segment.  residual, divergence_inf: per
 * cavity (max over its columns, the reference's definitions; in the mixed form those of the
 * last FP64 defect; iterations then sum the inner CG iterations over the sweeps).  Errors
 * as mac_stokes_solve;
 * the message names the first failing cavity. */
int32_t mac_stokes_solve_batched(void* handle, const double* rhs, const double* mass, const double* viscosity,
                                 double* velocity, double* pressure, int32_t* iterations, double* residual,
                                 double* divergence_inf, int32_t device_pointers);

#ifdef __cplusplus
}
#endif

#endif /* MACSTOKES_H */
