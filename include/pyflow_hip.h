/*
 * pyflow_hip.h -- C ABI of libpyflow_hip.so, the gfx950 HIP implementation of
 * PyFlow's projection time step.
 *
 * Conventions
 *   - Every function returns an int status: PF_OK (0) or a negative PF_E* code;
 *     pf_last_error() returns a thread-local message for the last failure.
 *   - Non-convergence and NaN/Inf are reported in output structs, never as
 *     errors (reference: linear_solvers/interface.py:51-53, solvers/solver.py:116-121).
 *   - Device buffers are owned by the caller: pf_batch_create() carves every
 *     vector out of one caller-provided device workspace and never frees it.
 *   - Host field arrays use the reference layout: float64, row-major
 *     (ny+2) x (nx+2) including the ghost ring (core/ghost_fields.py:47-60),
 *     or (ny) x (nx) interiors when ghost == 0; `ld` is the row stride in
 *     elements. Flat cell index k = j*nx + i (solvers/pressure_solver.py:19).
 *   - One batch = a set of independent cavities of arbitrary sizes advanced
 *     together; a single cavity is a batch of one.
 */
#ifndef PYFLOW_HIP_H
#define PYFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 2

#define PF_OK 0
#define PF_EINVAL (-1)
#define PF_ENOMEM (-2)
#define PF_EHIP (-3)
#define PF_ESTATE (-4)

/* Arithmetic of the N-vectors. FP64: every vector and stencil in float64
 * (parity form). MIXED: fields, stencils, SpMV and axpy in float32; dot
 * products, norms, the right-hand side and an outer iterative-refinement
 * residual/solution in float64. */
enum { PF_PREC_FP64 = 0, PF_PREC_MIXED = 1 };

/* Advection scheme (reference: solvers/solver.py:75-86). */
enum { PF_SCHEME_UPWIND = 0, PF_SCHEME_QUICK = 1, PF_SCHEME_NONE = 2 };

/* Pressure operator (reference: solvers/pressure_solver.py:61-106). */
enum { PF_OP_NEG_LAP_PINNED = 0, PF_OP_POS_LAP_DIRICHLET = 1 };

/* Standalone float64 operators (pf_op_apply). */
enum {
  PF_OPK_GRAD_COHERENT = 0,   /* in0=p          -> out0=gx, out1=gy  (numerics/core_ops_coherent.py:45-61) */
  PF_OPK_DIV_COHERENT = 1,    /* in0=u,in1=v    -> out0              (core_ops_coherent.py:63-79) */
  PF_OPK_LAP_COHERENT = 2,    /* in0=p          -> out0              (core_ops_coherent.py:81-84) */
  PF_OPK_DIV_CENTRAL = 3,     /* in0=u,in1=v    -> out0              (numerics/fluid_ops.py:8-28) */
  PF_OPK_ADV_QUICK = 4,       /* in0=u,in1=v,in2=f -> out0           (numerics/operators/advection.py:45-84) */
  PF_OPK_ADV_UPWIND = 5,      /* in0=u,in1=v,in2=f -> out0           (advection.py:17-43) */
  PF_OPK_NEG_LAP_PINNED = 6,  /* in0=x          -> out0              (pressure_solver.py:97-105) */
  PF_OPK_POS_LAP_DIRICHLET = 7/* in0=x          -> out0              (pressure_solver.py:69-77) */
};

/* What a batch run executes for each cavity. */
enum {
  PF_MODE_STEP = 0,     /* full step(): dt control, predictor, projection, lid BC, residuals */
  PF_MODE_PROJECT = 1,  /* solve_pressure_poisson() normal mode with a given dt */
  PF_MODE_SOLVE = 2     /* linear_solvers.solve(A, b, 'cg') on a PyFlow operator, x0 = 0 */
};

/* Per-cavity parameters; mirrors the attributes step() reads through getattr
 * (reference: solvers/solver.py:16-183, solvers/pressure_solver.py:58-146). */
typedef struct pf_cavity_params {
  int32_t nx, ny;          /* interior cells; both >= 2 */
  int32_t scheme;          /* PF_SCHEME_* */
  int32_t lid_enabled;     /* lid_velocity is not None */
  int32_t jacobi;          /* enable_jacobi_pc */
  int32_t lin_maxiter;
  int32_t op;              /* PF_OP_* */
  int32_t mode;            /* PF_MODE_* */
  double lid_velocity;
  double dx, dy;           /* lx/(nx-1), ly/(ny-1) */
  double nu;               /* 1/Re */
  double cfl_target, cfl_growth;
  double lin_tol;
  double inv_diag;         /* 1/(2/dx**2 + 2/dy**2): interior Jacobi weight */
  double inner_rtol;       /* MIXED only: relative tolerance of each FP32 inner PCG pass */
  /* Tiling and schedule; 0 = the library's default. pyflow.autotune times
   * the candidates on the device and picks them. Changing them changes the
   * order of the FP64 partial sums, not the arithmetic of any cell. */
  int32_t tile_rows;       /* region height in 16-row tiles: 1, 2, 4 or 8 (regions are 64 cells wide) */
  int32_t max_workgroups;  /* cap on the cavity's workgroups (regions are grouped to fit) */
  int32_t pcg_per_round;   /* MIXED: PCG iterations per round, 1 or 2 (s-step, s = 2) */
  int32_t iterate64;       /* MIXED: PCG iterate in 1 = FP32, 2 = FP64; 0 = FP64 from 1536 cells a side */
  /* MIXED, PF_OP_NEG_LAP_PINNED: after each solve, a Galerkin correction
   * x += y (y.r)/(y.A y) along y = A^-1 1, which removes the near-constant
   * error that FP32 rounding leaves along the pinned operator's slowest
   * mode. y is solved once per cavity (tolerance 1e-2) before its first solve.
   * 1 = off, 2 = on; 0 = on in PF_MODE_STEP and PF_MODE_PROJECT, off in
   * PF_MODE_SOLVE (a one-off solve would pay the set-up solve each time).
   * Ignored in FP64, which follows the reference's recurrence. */
  int32_t deflate;
  int32_t reserved;        /* 0 */
} pf_cavity_params;

/* Per-cavity results of the last completed step / projection / solve. */
typedef struct pf_step_diag {
  double dt, cfl, umax, vmax;      /* dt controller outputs (solver.py:25-62) */
  double dt_prev;                  /* == dt after a step; persisted for restart */
  double hist[5];                  /* divergence_history (last <=5 values) */
  int32_t hist_len;
  int32_t safeguard;               /* dt halved by the divergence-rise safeguard */
  double div_before;               /* ||central_div(u,v;1,1)|| at step start */
  double div_initial;              /* ||div_c(u*)|| before projection */
  double div_norm;                 /* ||div_c(u)|| after correction */
  double rhs_mean;                 /* mean subtracted from the Poisson RHS */
  double bnorm;                    /* ||b|| of the solved system */
  double rp_residual;              /* true ||b - A x|| */
  double ru, rv, continuity;       /* solver.py:108-114 */
  int32_t rp_info;                 /* SciPy info: 0 or lin_maxiter */
  int32_t rp_iterations;           /* PCG iterations actually executed (sum over passes) */
  int32_t rp_converged;            /* info == 0 and rnorm <= tol*||b|| */
  int32_t nan_detected;            /* any non-finite u, v, p */
  int32_t outer_passes;            /* MIXED: iterative-refinement passes */
  int32_t steps_done;              /* steps completed by the last pf_batch_run */
  int32_t rp_iterations_run;       /* PCG iterations summed over all steps of the last run */
  int32_t outer_passes_run;        /* refinement passes summed over the last run */
} pf_step_diag;

typedef struct pf_batch pf_batch;

int pf_abi_version(void);
const char* pf_last_error(void);
/* sizeof(pf_cavity_params) and sizeof(pf_step_diag), for binding layout checks. */
size_t pf_sizeof_params(void);
size_t pf_sizeof_diag(void);

/* Select the HIP device for the calling thread (hipSetDevice). */
int pf_device_init(int device);

/* Bytes of device workspace a batch needs. */
int pf_batch_workspace_bytes(int precision, int ncav, const pf_cavity_params* params, size_t* bytes);

/* Create a batch over `ncav` cavities. `workspace` is a device allocation of at
 * least pf_batch_workspace_bytes() bytes (256-byte aligned); `stream` is a
 * hipStream_t (NULL = default stream). Fields start at zero, dt_prev at 1e-2. */
int pf_batch_create(int precision, int ncav, const pf_cavity_params* params,
                    void* workspace, size_t workspace_bytes, void* stream, pf_batch** out);
int pf_batch_destroy(pf_batch* b);

/* Dynamic batches (continuous batching): an empty batch with room for
 * `max_cav` cavities and `max_tiles` workgroup regions in total; cavities are
 * added with their own workspace (pf_batch_workspace_bytes for ncav = 1) and
 * removed between runs. Slots are the indices used by every per-cavity call. */
int pf_batch_create_empty(int precision, int max_cav, int64_t max_tiles, void* stream, pf_batch** out);
int pf_batch_add(pf_batch* b, const pf_cavity_params* params, void* workspace, size_t bytes, int* slot);
int pf_batch_remove(pf_batch* b, int slot);

/* Replace the parameters of cavity `cav` (same nx, ny). The tiling chosen when
 * the cavity was added is kept. */
int pf_batch_set_params(pf_batch* b, int cav, const pf_cavity_params* params);

/* The parameters cavity `cav` runs with, every library default resolved:
 * tile_rows = its region height, max_workgroups = its workgroup count,
 * pcg_per_round (1 or 2), iterate64 (1 or 2; 0 in FP64) and deflate (1 or 2;
 * 0 in FP64). Passing them back
 * (pf_batch_add / pf_batch_create) reproduces the same tiling and so the same
 * FP64 partial-sum order whatever PYFLOW_* environment the process has; the
 * checkpoint meta written at io/checkpoint.py:66-76 carries them so that a
 * restart continues bit for bit. */
int pf_batch_resolved_params(pf_batch* b, int cav, pf_cavity_params* out);

/* Host float64 <-> device fields of one cavity. Any of u, v, p may be NULL.
 * Called between runs. MIXED fields are converted through an FP64 staging area
 * inside the cavity's own workspace (the predictor buffers, which hold nothing
 * between steps), so the right-hand side of pf_batch_set_rhs and the last
 * solution stay intact. */
int pf_batch_upload(pf_batch* b, int cav, const double* u, const double* v, const double* p,
                    int64_t ld, int ghost);
int pf_batch_download(pf_batch* b, int cav, double* u, double* v, double* p,
                      int64_t ld, int ghost);
/* Zero u, v, p of every cavity (allocate_state()). */
int pf_batch_zero_fields(pf_batch* b);

/* dt controller memory: state.meta['dt_prev'] and ['divergence_history']. */
int pf_batch_set_meta(pf_batch* b, int cav, double dt_prev, const double* hist, int hist_len);

/* Right-hand side for PF_MODE_SOLVE (float64, ny x nx, row stride ld). */
int pf_batch_set_rhs(pf_batch* b, int cav, const double* rhs, int64_t ld);
/* Solution of the last solve/projection in float64 (x before any pin offset). */
int pf_batch_get_solution(pf_batch* b, int cav, double* x, int64_t ld);

/* dt used by PF_MODE_PROJECT. */
int pf_batch_set_dt(pf_batch* b, int cav, double dt);

/* Run: every cavity whose `active` flag is non-zero (NULL = all) executes
 * `nsteps` steps (PF_MODE_STEP) or one projection/solve; blocks until all are
 * done, then writes one pf_step_diag per cavity into `diag` (may be NULL). */
int pf_batch_run(pf_batch* b, int nsteps, const int32_t* active, pf_step_diag* diag);

/* Start the same work without waiting (kernels queued on the batch stream
 * in chunks); pf_batch_poll() returns 1 when finished, 0 if not yet. */
int pf_batch_launch(pf_batch* b, int nsteps, const int32_t* active);
int pf_batch_poll(pf_batch* b, int* done);
int pf_batch_diag(pf_batch* b, pf_step_diag* diag);

/* Continuous batching (served solves; no counterpart in the reference, whose
 * driver advances one cavity per process, drivers/simulation_driver.py:58-110).
 * Between chunks of rounds:
 *   pf_batch_start    arms cavity `slot` for `nsteps` steps (one projection or
 *                     solve in the other modes); the other cavities keep their
 *                     device-resident state and continue where they are;
 *   pf_batch_advance  queues `rounds` rounds over the batch and returns at once;
 *                     pf_batch_poll() then reports 1 when they have run and
 *                     queues nothing more;
 *   pf_batch_finished writes finished[slot] = 1 for every cavity that completed
 *                     its work (then pf_batch_diag has its results and the slot
 *                     may be removed). `finished` has max_cav entries. */
int pf_batch_start(pf_batch* b, int slot, int nsteps);
int pf_batch_advance(pf_batch* b, int64_t rounds);
int pf_batch_finished(pf_batch* b, int32_t* finished);

/* u[:, nx//2] of one cavity as float64 (drivers/simulation_driver.py:90-96). */
int pf_batch_centerline(pf_batch* b, int cav, double* out);

/* Rounds (kernel launches) executed by the last run; for profiling. */
int pf_batch_rounds(pf_batch* b, int64_t* rounds);
/* Optional timing: with timing enabled every chunk of back-to-back round
 * launches is bracketed by hipEvents on the batch stream; pf_batch_timing
 * returns the summed GPU milliseconds and the number of rounds they cover
 * (host-side polling gaps excluded). Enabling resets both counters. */
int pf_batch_set_timing(pf_batch* b, int enable);
int pf_batch_timing(pf_batch* b, double* gpu_ms, int64_t* rounds);

/* Rounds are queued as replays of an instantiated hipGraph: a chain of
 * `rounds` kernel nodes (rounded down to an even count, at most 1024; 0
 * queues every round with its own launch). Replays and direct launches run
 * the same kernels with the same arguments, so results are bitwise equal.
 * The graphs are rebuilt after any change of the batch's cavities. Default:
 * PYFLOW_GRAPH if set, else 64. The reference has no counterpart (its step
 * runs on the host: solvers/solver.py:15); this serves the batched and
 * served paths (drivers/simulation_driver.py:58-110 per cavity). */
int pf_batch_set_graph(pf_batch* b, int rounds);

/* Standalone float64 operator on device arrays of shape (ny, nx), dense rows.
 * Unused inputs/outputs may be NULL. */
int pf_op_apply(int opk, int nx, int ny, double dx, double dy,
                const double* in0, const double* in1, const double* in2,
                double* out0, double* out1, void* stream);

/* Allocation-free solve. No counterpart in the reference: it serves callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid hipMalloc). It is the PF_MODE_SOLVE run of pf_batch_create +
 * pf_batch_set_rhs + pf_batch_run + pf_batch_get_solution on one cavity
 * (linear_solvers.solve on a PyFlow operator, reference
 * linear_solvers/interface.py:23-53): the same kernels and tiling, so the same
 * bits. `rhs` (read only) and `x` (the FP64 solution, written) are device
 * arrays of ny x nx float64 in dense rows; every other buffer lives in a
 * static device array of PF_NOALLOC_SCRATCH_BYTES, one per GPU, held until
 * the process ends. A cavity whose buffers do not fit returns PF_EINVAL. No
 * hipMalloc, hipFree, event or graph is created. Blocks until the solve is
 * done (it polls its stream, so it cannot be captured into a graph). Calls
 * that overlap on one GPU race on that GPU's scratch array. `diag` may be
 * NULL. */
#define PF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int pf_solve_noalloc(int precision, const pf_cavity_params* params, const double* rhs, double* x,
                     void* stream, pf_step_diag* diag);

#ifdef __cplusplus
}
#endif
#endif /* PYFLOW_HIP_H */
