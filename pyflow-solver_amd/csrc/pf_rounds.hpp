// Device phase machine of the projection step: per-cell phase bodies, the
// scalar "tails" that consume a phase's reduced partials, and the round kernel.
//
// Cross-workgroup protocol (no fences, no spin-waits)
// -------------------------------------------------
// Cavity c is covered by nblk workgroups; workgroup b owns tiles
// [b*tpb, (b+1)*tpb) of 64x16 cells (one cell per thread per tile). In round r
// every workgroup of c
//   1. copies the cavity's scalar state written in round r-1 (buffer r^1),
//   2. if that state has a pending phase, sums the pending phase's FP64
//      partials of all nblk workgroups (written in round r-1, parity r^1) in
//      a fixed order and runs the phase's scalar tail (dt controller, CG
//      alpha/beta, convergence tests, phase transition) -- every workgroup
//      computes the same bits, so they all agree on the next phase,
//   3. executes that phase on its tiles and writes its own partials (parity r),
//   4. (workgroup 0 only) stores the state, with the executed phase pending,
//      into buffer r.
// Data written in round r is only read in round r+1, so kernel boundaries are
// the only synchronisation, and all reductions are bitwise deterministic.
#pragma once
#include <hip/hip_runtime.h>

#include "pf_cells.hpp"
#include "pyflow_hip.h"

namespace pf {

constexpr int TX = 64;        // tile width = wavefront width
constexpr int TY = 16;        // tile height; one wave per tile row
constexpr int NTH = TX * TY;  // threads per workgroup
constexpr int NWAVE = NTH / 64;
constexpr int NP = 4;         // FP64 partial slots per workgroup

enum Phase {
  PH_IDLE = 0,
  PH_STATS = 1,
  PH_PRED_A = 2,
  PH_PRED_B = 3,
  PH_PINIT = 4,
  PH_P = 5,
  PH_Q = 6,
  PH_OUTER_A = 7,
  PH_OUTER_B = 8,
  PH_FIN = 9,
  PH_CORR_A = 10,
  PH_CORR_B = 11,
  PH_CG0 = 12,  // MIXED: w = A M r and the first dot products of an inner solve
  PH_CG = 13,   // MIXED: one single-reduction (Chronopoulos-Gear) PCG iteration
  PH_DONE = 14,
};

// Partials per phase and which of them are NaN-propagating maxima.
__device__ __host__ inline int phase_nparts(int ph) {
  switch (ph) {
    case PH_STATS: return 3;
    case PH_PRED_B: return 2;
    case PH_PINIT: return 3;
    case PH_P: return 1;
    case PH_Q: return 2;
    case PH_OUTER_B: return 3;
    case PH_FIN: return 1;
    case PH_CORR_A: return 3;
    case PH_CORR_B: return 2;
    case PH_CG0: return 3;
    case PH_CG: return 3;
    default: return 0;
  }
}
__device__ __host__ inline unsigned phase_maxmask(int ph) {
  return ph == PH_STATS ? 0x3u : (ph == PH_CORR_A ? 0x4u : 0u);
}

// Read-mostly description of one cavity (rewritten only between runs).
struct PfCav {
  int nx, ny, tiles_x, tiles_y, ntiles, nblk, blk0, tpb;  // tpb = M: tiles per workgroup, stacked in y
  int scheme, lid_on, jacobi, maxiter, op, mode, prec;
  double lid, dx, dy, nu, cfl, growth, tol, invd, inner_rtol;
  void *u, *v, *p, *us, *vs, *r, *d0, *d1, *x;  // element type T
  void *r1, *w0, *w1, *pd;                      // MIXED single-reduction PCG: r, w = A M r (two
                                                // buffers each), search direction; s = A p in d0/d1
  double *b, *x64;                              // x64 aliases x in FP64 mode
  double* part;                                 // [2][nblk][NP]
};

// Scalar state of one cavity; two copies (round parity) per cavity.
struct PfScal {
  int phase;     // phase to execute next
  int pending;   // phase whose partials await their tail (0 = none)
  int steps_left, steps_done, it, it_total, info, outer, x_is_b;
  int hist_len, safeguard, nan, converged, it_run, outer_run, done_now;
  double dt_prev, dt, cfl_now, umax, vmax, hist[5], div_before;
  double mean, div_initial, bnorm, atol, inner_atol, rho, rho_prev, beta, pq, alpha, rnorm2, rn_outer;
  double rnorm_true, div_after, ru, rv, continuity;
};
static_assert(sizeof(PfScal) % 4 == 0, "PfScal copied as 32-bit words");

template <class T>
struct Grid {
  const T* __restrict__ a;
  int nx;
  __device__ T operator()(int j, int i) const { return a[(size_t)j * nx + i]; }
};

__device__ inline int tid_flat() { return threadIdx.y * TX + threadIdx.x; }

__device__ inline double redop(unsigned maxmask, int k, double a, double b) {
  return ((maxmask >> k) & 1u) ? nanmax(a, b) : a + b;
}

// Wave reduction on the DPP crossbar (no LDS): quad swaps, row half-mirror and
// row mirror give 16-lane totals in every lane, then row_bcast15/31 fold the
// four rows into lane 63. The combination order is fixed, so the result is
// bitwise reproducible and identical in every workgroup reducing equal data.
// Rows/lanes outside the DPP row mask read 0, the identity of both the sum and
// the maximum of non-negative values used here.
template <bool MAX>
__device__ inline double red2(double a, double b) {
  if constexpr (MAX)
    return nanmax(a, b);
  else
    return a + b;
}

template <int CTRL, int ROWMASK, bool MAX>
__device__ inline double dpp_step(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, ROWMASK, 0xF, false);
  const long long ob = (long long)(unsigned)lo | ((long long)hi << 32);
  return red2<MAX>(v, __longlong_as_double(ob));
}

template <bool MAX>
__device__ inline double wave_reduce(double v) {
  v = dpp_step<0xB1, 0xF, MAX>(v);   // quad_perm [1,0,3,2]
  v = dpp_step<0x4E, 0xF, MAX>(v);   // quad_perm [2,3,0,1]
  v = dpp_step<0x141, 0xF, MAX>(v);  // row_half_mirror
  v = dpp_step<0x140, 0xF, MAX>(v);  // row_mirror
  v = dpp_step<0x142, 0xA, MAX>(v);  // row_bcast15 into rows 1, 3
  v = dpp_step<0x143, 0xC, MAX>(v);  // row_bcast31 into rows 2, 3
  return v;                          // total in lane 63
}

// Workgroup reduction of the first K of NP per-thread values (bit k of MM set:
// NaN-propagating maximum, else sum) into out[0..K) in shared memory. Ends
// with a barrier, so every thread may read out[] afterwards.
template <int K, unsigned MM>
__device__ inline void block_reduce_k(double (&v)[NP], double (*sm)[NP], double* out) {
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((MM >> k) & 1u) ? wave_reduce<true>(v[k]) : wave_reduce<false>(v[k]);
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[wave][k] = v[k];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = (lane < NWAVE) ? sm[lane][k] : 0.0;
      t = ((MM >> k) & 1u) ? wave_reduce<true>(t) : wave_reduce<false>(t);
      if (lane == 63) out[k] = t;
    }
  }
  __syncthreads();
}

// Dispatch on a phase's partial layout (compile-time K and reduction ops).
__device__ inline void block_reduce_phase(int ph, double (&v)[NP], double (*sm)[NP], double* out) {
  switch (ph) {
    case PH_STATS: block_reduce_k<3, 0x3u>(v, sm, out); break;
    case PH_CORR_A: block_reduce_k<3, 0x4u>(v, sm, out); break;
    case PH_PINIT:
    case PH_OUTER_B:
    case PH_CG0:
    case PH_CG: block_reduce_k<3, 0u>(v, sm, out); break;
    case PH_PRED_B:
    case PH_Q:
    case PH_CORR_B: block_reduce_k<2, 0u>(v, sm, out); break;
    case PH_P:
    case PH_FIN: block_reduce_k<1, 0u>(v, sm, out); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Scalar tails: run by one thread on the workgroup's shared copy of the state
// after the pending phase's partials have been summed into tot[].
// ---------------------------------------------------------------------------
__device__ inline void cg_next(const PfCav& C, PfScal& S) {
  const bool mixed = (C.prec == PF_PREC_MIXED);
  if (S.it_total >= C.maxiter) {
    S.info = C.maxiter;
    S.phase = mixed ? PH_OUTER_A : PH_FIN;
    return;
  }
  if (sqrt(S.rnorm2) < S.inner_atol) {
    if (!mixed) S.info = 0;
    S.phase = mixed ? PH_OUTER_A : PH_FIN;
    return;
  }
  S.phase = mixed ? PH_CG : PH_P;
}

__device__ inline void enter_done(PfScal& S) {
  S.phase = PH_DONE;
  S.done_now = 1;
}

__device__ void tail(const PfCav& C, PfScal& S, const double* tot) {
  const int ph = S.pending;
  S.pending = PH_IDLE;
  switch (ph) {
    case PH_STATS: {  // dt controller, solver.py:25-62 (Python float semantics)
      const double umax = tot[0] + 1e-12;
      const double vmax = tot[1] + 1e-12;
      const double dt_cfl = pymin(C.cfl * C.dx / umax, C.cfl * C.dy / vmax);
      const double dt_diff = 0.25 * pymin(C.dx * C.dx, C.dy * C.dy) / pymax(C.nu, 1e-12);
      const double dt_cand = pymin(dt_cfl, dt_diff);
      const double cap = S.dt_prev * C.growth;
      double dt = (dt_cand > cap) ? cap : dt_cand;
      const double div0 = sqrt(tot[2]);
      int hl = S.hist_len;
      if (hl == 5) {
        for (int k = 0; k < 4; ++k) S.hist[k] = S.hist[k + 1];
        S.hist[4] = div0;
      } else {
        S.hist[hl] = div0;
        hl += 1;
        S.hist_len = hl;
      }
      int sg = 0;
      if (hl >= 3) {
        const double a = S.hist[hl - 3], b = S.hist[hl - 2], c = S.hist[hl - 1];
        if (a > 0 && b > a * 1.01 && c > b * 1.01) {
          dt *= 0.5;
          sg = 1;
        }
      }
      S.safeguard = sg;
      S.dt = dt;
      S.dt_prev = dt;
      S.umax = umax;
      S.vmax = vmax;
      S.div_before = div0;
      S.cfl_now = pymax(umax * dt / C.dx, vmax * dt / C.dy);
      S.phase = PH_PRED_A;
      break;
    }
    case PH_PRED_A: S.phase = PH_PRED_B; break;
    case PH_PRED_B:
      S.mean = tot[0] / (double)((size_t)C.nx * C.ny);
      S.div_initial = sqrt(tot[1]);
      S.phase = PH_PINIT;
      break;
    case PH_PINIT: {
      const double bn = sqrt(tot[0]);
      S.bnorm = bn;
      S.atol = pymax(0.0, C.tol * bn);
      S.it = 0;
      S.it_total = 0;
      S.outer = 0;
      S.info = 0;
      S.x_is_b = 0;
      if (bn == 0.0) {  // SciPy returns x = b with info 0
        S.x_is_b = 1;
        S.phase = PH_FIN;
        break;
      }
      S.rho = tot[1];
      S.rnorm2 = tot[2];
      S.rn_outer = bn;
      if (C.maxiter <= 0) {  // the loop never runs: info = maxiter
        S.info = C.maxiter;
        S.phase = PH_FIN;
        break;
      }
      if (C.prec == PF_PREC_MIXED) {
        if (bn < S.atol) {
          S.phase = PH_FIN;
          break;
        }
        S.inner_atol = pymax(C.inner_rtol * bn, 0.5 * S.atol);
        S.phase = PH_CG0;
        break;
      }
      S.inner_atol = S.atol;
      S.phase = (sqrt(S.rnorm2) < S.atol) ? PH_FIN : PH_P;
      break;
    }
    case PH_P:
      S.pq = tot[0];
      S.alpha = S.rho / tot[0];
      S.phase = PH_Q;
      break;
    case PH_Q:
      S.rho_prev = S.rho;
      S.rho = tot[0];
      S.rnorm2 = tot[1];
      S.beta = S.rho / S.rho_prev;
      S.it += 1;
      S.it_total += 1;
      S.it_run += 1;
      cg_next(C, S);
      break;
    case PH_OUTER_A: S.phase = PH_OUTER_B; break;
    case PH_OUTER_B: {
      const double rn = sqrt(tot[0]);
      S.rn_outer = rn;
      S.outer += 1;
      S.outer_run += 1;
      if (S.info != 0) {  // iteration budget exhausted
        S.phase = PH_FIN;
        break;
      }
      if (rn < S.atol) {
        S.info = 0;
        S.phase = PH_FIN;
        break;
      }
      S.it = 0;
      S.inner_atol = pymax(C.inner_rtol * rn, 0.5 * S.atol);
      S.phase = PH_CG0;
      break;
    }
    case PH_CG0:  // gamma0 = r.Mr, delta0 = (A M r).(M r), alpha0 = gamma0/delta0
      S.rho = tot[0];
      S.alpha = tot[0] / tot[1];
      S.beta = 0.0;
      S.rnorm2 = tot[2];
      S.it = 0;
      if (S.it_total >= C.maxiter) {
        S.info = C.maxiter;
        S.phase = PH_OUTER_A;
      } else {
        S.phase = (sqrt(S.rnorm2) < S.inner_atol) ? PH_OUTER_A : PH_CG;
      }
      break;
    case PH_CG: {  // beta = gamma'/gamma, alpha' = gamma'/(delta' - beta gamma'/alpha)
      const double g1 = tot[0], d1 = tot[1];
      const double beta = g1 / S.rho;
      S.alpha = g1 / (d1 - beta * g1 / S.alpha);
      S.beta = beta;
      S.rho = g1;
      S.rnorm2 = tot[2];
      S.it += 1;
      S.it_total += 1;
      S.it_run += 1;
      cg_next(C, S);
      break;
    }
    case PH_FIN: {
      const double rn = sqrt(tot[0]);
      S.rnorm_true = rn;
      const double bn = (S.bnorm == 0.0) ? 1.0 : S.bnorm;
      S.converged = (S.info == 0) && (rn <= C.tol * bn);
      if (C.mode == PF_MODE_SOLVE) {
        S.steps_done += 1;
        enter_done(S);
      } else {
        S.phase = PH_CORR_A;
      }
      break;
    }
    case PH_CORR_A: {
      const double den = pymax(1.0, (double)((size_t)C.nx * C.ny));
      S.ru = sqrt(tot[0]) / den;
      S.rv = sqrt(tot[1]) / den;
      S.nan = (tot[2] != 0.0) ? 1 : 0;
      S.phase = PH_CORR_B;
      break;
    }
    case PH_CORR_B:
      S.div_after = sqrt(tot[0]);
      S.continuity = sqrt(tot[1]) / pymax(1.0, (double)((size_t)C.nx * C.ny));
      S.steps_done += 1;
      S.steps_left -= 1;
      if (C.mode == PF_MODE_STEP && S.steps_left > 0)
        S.phase = PH_STATS;
      else
        enter_done(S);
      break;
    default: break;
  }
}

// Scalars a phase body reads, hoisted out of the shared state.
struct PhaseArgs {
  double dt, mean, beta, alpha, x0;
  int it, x_is_b;
  float cx, cy, betaf, alphaf, minvf;  // float32 path: 1/dx^2, 1/dy^2, CG scalars, Jacobi weight
};

// ---------------------------------------------------------------------------
// Per-cell phase bodies; acc[] accumulates this thread's partials.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline void cell_stats(const PfCav& C, int j, int i, double* acc) {
  Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
  acc[0] = nanmax(acc[0], fabs((double)U(j, i)));
  acc[1] = nanmax(acc[1], fabs((double)V(j, i)));
  const double cd = (double)div_central<T>(U, V, C.nx, C.ny, j, i, T(1), T(1));
  acc[2] += cd * cd;
}

template <class T>
__device__ inline void cell_pred_a(const PfCav& C, const PhaseArgs& A, int j, int i) {
  Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
  const size_t k = (size_t)j * C.nx + i;
  const T uc = U(j, i), vc = V(j, i);
  T us = uc, vs = vc;
  if (C.scheme != PF_SCHEME_NONE) {
    const T dt = (T)A.dt, dx = (T)C.dx, dy = (T)C.dy;
    T cu, cv;
    if (C.scheme == PF_SCHEME_QUICK) {
      cu = adv_quick<T>(uc, vc, U, C.nx, C.ny, j, i, dx, dy);
      cv = adv_quick<T>(uc, vc, V, C.nx, C.ny, j, i, dx, dy);
    } else {
      cu = adv_upwind<T>(uc, vc, U, C.nx, C.ny, j, i, dx, dy);
      cv = adv_upwind<T>(uc, vc, V, C.nx, C.ny, j, i, dx, dy);
    }
    us = uc - dt * cu;
    vs = vc - dt * cv;
  }
  ((T*)C.us)[k] = us;
  ((T*)C.vs)[k] = vs;
}

template <class T>
__device__ inline void cell_pred_b(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  Grid<T> US{(const T*)C.us, C.nx}, VS{(const T*)C.vs, C.nx};
  const T dv = div_coherent<T>(US, VS, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
  const double rhs = (-(double)dv) / pymax(A.dt, 1e-14);
  C.b[(size_t)j * C.nx + i] = rhs;
  acc[0] += rhs;
  acc[1] += (double)dv * (double)dv;
}

template <class T>
__device__ inline void cell_pinit(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const size_t k = (size_t)j * C.nx + i;
  double bb = C.b[k];
  if (C.mode != PF_MODE_SOLVE) {
    bb = bb - A.mean;
    if (k == 0) bb = 0.0;
    C.b[k] = bb;
  }
  const T rr = (T)bb;
  ((T*)C.r)[k] = rr;
  ((T*)C.x)[k] = T(0);
  if (C.prec == PF_PREC_MIXED) C.x64[k] = 0.0;
  const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rr;
  acc[0] += bb * bb;
  acc[1] += (double)rr * (double)z;
  acc[2] += (double)rr * (double)rr;
}

// Search direction: d = z on the first inner iteration, else (d_old*beta) + z.
template <class T>
__device__ inline T dnew_at(const PfCav& C, const T* __restrict__ R, const T* __restrict__ Dold, int first,
                            double beta, int j, int i) {
  const size_t k = (size_t)j * C.nx + i;
  const T rr = R[k];
  const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rr;
  if (first) return z;
  return (T)((double)Dold[k] * beta + (double)z);
}

template <class T>
__device__ inline void cell_p(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const T* R = (const T*)C.r;
  const T* Dold = (const T*)((A.it & 1) ? C.d0 : C.d1);
  T* Dnew = (T*)((A.it & 1) ? C.d1 : C.d0);
  const int first = (A.it == 0);
  const int nx = C.nx, ny = C.ny;
  const T c = dnew_at<T>(C, R, Dold, first, A.beta, j, i);
  const T e = (i + 1 < nx) ? dnew_at<T>(C, R, Dold, first, A.beta, j, i + 1) : T(0);
  const T w = (i > 0) ? dnew_at<T>(C, R, Dold, first, A.beta, j, i - 1) : T(0);
  const T n = (j + 1 < ny) ? dnew_at<T>(C, R, Dold, first, A.beta, j + 1, i) : T(0);
  const T s = (j > 0) ? dnew_at<T>(C, R, Dold, first, A.beta, j - 1, i) : T(0);
  Dnew[(size_t)j * nx + i] = c;
  const T q = apply_op5<T>(C.op, c, e, w, n, s, nx, ny, j, i, (T)C.dx, (T)C.dy);
  acc[0] += (double)c * (double)q;
}

template <class T>
__device__ inline void cell_q(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const T* D = (const T*)((A.it & 1) ? C.d1 : C.d0);
  Grid<T> DG{D, C.nx};
  const size_t k = (size_t)j * C.nx + i;
  const T q = apply_op_at<T>(C.op, DG, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
  T* X = (T*)C.x;
  T* R = (T*)C.r;
  const T d = D[k];
  X[k] = (T)((double)X[k] + A.alpha * (double)d);
  const T rn = (T)((double)R[k] - A.alpha * (double)q);
  R[k] = rn;
  const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rn;
  acc[0] += (double)rn * (double)z;
  acc[1] += (double)rn * (double)rn;
}

// Float32 helpers of the mixed-precision P/Q: Jacobi weight and
// d = fma(d_old, beta, z) in float32.
__device__ inline float minv_f32(const PfCav& C, const PhaseArgs& A, int j, int i) {
  if (!C.jacobi) return 1.0f;
  const bool bnd = (i == 0 || i == C.nx - 1 || j == 0 || j == C.ny - 1);
  return bnd ? 1.0f : A.minvf;
}
__device__ inline float dnew_f32(const PfCav& C, const PhaseArgs& A, const float* __restrict__ R,
                                 const float* __restrict__ Dold, int first, int j, int i) {
  const size_t k = (size_t)j * C.nx + i;
  const float z = minv_f32(C, A, j, i) * R[k];
  if (first) return z;
  return fmaf(Dold[k], A.betaf, z);
}

// P and Q over a chunk of U tiles: all loads of the chunk are issued before
// any store (stores go to other arrays than the loads, but the compiler cannot
// prove it through the cavity descriptor), so a thread keeps U cells in flight.
// Workgroup b of a cavity owns the region of M = tpb tiles stacked in y:
// tile column b % tiles_x, tile rows (b / tiles_x) * M + m, m = 0..M-1.
struct TileIt {
  int tx, ty, tyend;
  __device__ void next(int) { ++ty; }
};

template <class T, int U>
__device__ inline void chunk_p(const PfCav& C, const PhaseArgs& A, TileIt& ti, double* acc) {
  const T* R = (const T*)C.r;
  const T* Dold = (const T*)((A.it & 1) ? C.d0 : C.d1);
  T* Dnew = (T*)((A.it & 1) ? C.d1 : C.d0);
  const int first = (A.it == 0);
  const int nx = C.nx, ny = C.ny;
  T cv[U], qv[U];
  size_t kk[U];
  bool ok[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = ti.tx * TX + threadIdx.x;
    const int j = ti.ty * TY + threadIdx.y;
    ok[u] = (ti.ty < ti.tyend) && (i < nx) && (j < ny);
    kk[u] = (size_t)j * nx + i;
    ti.next(C.tiles_x);
    cv[u] = T(0);
    qv[u] = T(0);
    if (ok[u]) {
      if constexpr (sizeof(T) == sizeof(float)) {
        const float c = dnew_f32(C, A, (const float*)R, (const float*)Dold, first, j, i);
        const float e = (i + 1 < nx) ? dnew_f32(C, A, (const float*)R, (const float*)Dold, first, j, i + 1) : 0.f;
        const float w = (i > 0) ? dnew_f32(C, A, (const float*)R, (const float*)Dold, first, j, i - 1) : 0.f;
        const float n = (j + 1 < ny) ? dnew_f32(C, A, (const float*)R, (const float*)Dold, first, j + 1, i) : 0.f;
        const float s = (j > 0) ? dnew_f32(C, A, (const float*)R, (const float*)Dold, first, j - 1, i) : 0.f;
        cv[u] = c;
        qv[u] = apply_op5_fast(C.op, c, e, w, n, s, nx, ny, j, i, A.cx, A.cy);
      } else {
        const T c = dnew_at<T>(C, R, Dold, first, A.beta, j, i);
        const T e = (i + 1 < nx) ? dnew_at<T>(C, R, Dold, first, A.beta, j, i + 1) : T(0);
        const T w = (i > 0) ? dnew_at<T>(C, R, Dold, first, A.beta, j, i - 1) : T(0);
        const T n = (j + 1 < ny) ? dnew_at<T>(C, R, Dold, first, A.beta, j + 1, i) : T(0);
        const T s = (j > 0) ? dnew_at<T>(C, R, Dold, first, A.beta, j - 1, i) : T(0);
        cv[u] = c;
        qv[u] = apply_op5<T>(C.op, c, e, w, n, s, nx, ny, j, i, (T)C.dx, (T)C.dy);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (ok[u]) {
      Dnew[kk[u]] = cv[u];
      acc[0] += (double)cv[u] * (double)qv[u];
    }
  }
}

template <class T, int U>
__device__ inline void chunk_q(const PfCav& C, const PhaseArgs& A, TileIt& ti, double* acc) {
  const T* D = (const T*)((A.it & 1) ? C.d1 : C.d0);
  Grid<T> DG{D, C.nx};
  T* X = (T*)C.x;
  T* R = (T*)C.r;
  const int nx = C.nx, ny = C.ny;
  T xv[U], rv[U], zw[U];
  size_t kk[U];
  bool ok[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = ti.tx * TX + threadIdx.x;
    const int j = ti.ty * TY + threadIdx.y;
    ok[u] = (ti.ty < ti.tyend) && (i < nx) && (j < ny);
    kk[u] = (size_t)j * nx + i;
    ti.next(C.tiles_x);
    xv[u] = T(0);
    rv[u] = T(0);
    zw[u] = T(0);
    if (ok[u]) {
      if constexpr (sizeof(T) == sizeof(float)) {
        const float c = DG(j, i);
        const float e = (i + 1 < nx) ? DG(j, i + 1) : 0.f;
        const float w = (i > 0) ? DG(j, i - 1) : 0.f;
        const float n = (j + 1 < ny) ? DG(j + 1, i) : 0.f;
        const float s = (j > 0) ? DG(j - 1, i) : 0.f;
        const float q = apply_op5_fast(C.op, c, e, w, n, s, nx, ny, j, i, A.cx, A.cy);
        xv[u] = fmaf(A.alphaf, c, X[kk[u]]);
        rv[u] = fmaf(-A.alphaf, q, R[kk[u]]);
        zw[u] = minv_f32(C, A, j, i);
      } else {
        const T q = apply_op_at<T>(C.op, DG, nx, ny, j, i, (T)C.dx, (T)C.dy);
        const T d = D[kk[u]];
        xv[u] = (T)((double)X[kk[u]] + A.alpha * (double)d);
        rv[u] = (T)((double)R[kk[u]] - A.alpha * (double)q);
        zw[u] = (T)jacobi_minv(C.jacobi, C.invd, nx, ny, j, i);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (ok[u]) {
      X[kk[u]] = xv[u];
      R[kk[u]] = rv[u];
      const T z = zw[u] * rv[u];
      acc[0] += (double)rv[u] * (double)z;
      acc[1] += (double)rv[u] * (double)rv[u];
    }
  }
}

template <class T>
__device__ inline void cell_outer_a(const PfCav& C, int j, int i) {
  const size_t k = (size_t)j * C.nx + i;
  T* X = (T*)C.x;
  C.x64[k] = C.x64[k] + (double)X[k];
  X[k] = T(0);
}

template <class T>
__device__ inline void cell_outer_b(const PfCav& C, int j, int i, double* acc) {
  const size_t k = (size_t)j * C.nx + i;
  Grid<double> XG{C.x64, C.nx};
  const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
  const double r64 = C.b[k] - ax;
  const T rt = (T)r64;
  ((T*)C.r)[k] = rt;
  const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rt;
  acc[0] += r64 * r64;
  acc[1] += (double)rt * (double)z;
  acc[2] += (double)rt * (double)rt;
}

template <class T>
__device__ inline void cell_fin(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const size_t k = (size_t)j * C.nx + i;
  const double* XS = A.x_is_b ? C.b : C.x64;
  Grid<double> XG{XS, C.nx};
  const double xv = XS[k];
  const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
  const double rr = C.b[k] - ax;
  acc[0] += rr * rr;
  const double pv = (C.mode == PF_MODE_SOLVE) ? xv : (xv - A.x0);
  ((T*)C.p)[k] = (T)pv;
}

template <class T>
__device__ inline void cell_corr_a(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const int nx = C.nx, ny = C.ny;
  const size_t k = (size_t)j * nx + i;
  Grid<T> PG{(const T*)C.p, nx};
  const T dt = (T)A.dt;
  const T gx = grad_x_coherent<T>(PG, nx, ny, j, i, (T)C.dx);
  const T gy = grad_y_coherent<T>(PG, nx, ny, j, i, (T)C.dy);
  T* US = (T*)C.us;
  T* VS = (T*)C.vs;
  const T uc = US[k] - dt * gx;
  const T vc = VS[k] - dt * gy;
  US[k] = uc;
  VS[k] = vc;
  const T uf = (C.lid_on && j == ny - 1) ? (T)C.lid : uc;
  T* U = (T*)C.u;
  T* V = (T*)C.v;
  const double du = (double)uf - (double)U[k];
  const double dv = (double)vc - (double)V[k];
  U[k] = uf;
  V[k] = vc;
  acc[0] += du * du;
  acc[1] += dv * dv;
  const bool bad = !isfinite((double)uf) || !isfinite((double)vc) || !isfinite((double)PG(j, i));
  acc[2] = nanmax(acc[2], bad ? 1.0 : 0.0);
}

template <class T>
__device__ inline void cell_corr_b(const PfCav& C, int j, int i, double* acc) {
  Grid<T> US{(const T*)C.us, C.nx}, VS{(const T*)C.vs, C.nx};
  Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
  const double dc = (double)div_coherent<T>(US, VS, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
  const double cd = (double)div_central<T>(U, V, C.nx, C.ny, j, i, T(1), T(1));
  acc[0] += dc * dc;
  acc[1] += cd * cd;
}

// MIXED inner PCG, single reduction per iteration (Chronopoulos & Gear):
//   s_i = w_i + beta_i s_{i-1};  p_i = u_i + beta_i p_{i-1};  u = M r, w = A u
//   x_{i+1} = x_i + alpha_i p_i; r_{i+1} = r_i - alpha_i s_i
//   w_{i+1} = A M r_{i+1};  partials r.Mr, w.Mr, r.r of iteration i+1.
// r_{i+1} is recomputed at the four neighbours from r_i, w_i, s_{i-1}, so r,
// w and s are double-buffered by iteration parity; p and x are updated in place.
template <class T>
__device__ inline void cell_cg0(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const int nx = C.nx, ny = C.ny;
  const T* R = (const T*)C.r;
  auto U = [&](int jj, int ii) -> T {
    T m;
    if constexpr (sizeof(T) == sizeof(float)) m = minv_f32(C, A, jj, ii);
    else m = (T)jacobi_minv(C.jacobi, C.invd, nx, ny, jj, ii);
    return m * R[(size_t)jj * nx + ii];
  };
  const T c = U(j, i);
  const T e = (i + 1 < nx) ? U(j, i + 1) : T(0);
  const T w = (i > 0) ? U(j, i - 1) : T(0);
  const T n = (j + 1 < ny) ? U(j + 1, i) : T(0);
  const T s = (j > 0) ? U(j - 1, i) : T(0);
  T wv;
  if constexpr (sizeof(T) == sizeof(float)) wv = apply_op5_fast(C.op, c, e, w, n, s, nx, ny, j, i, A.cx, A.cy);
  else wv = apply_op5<T>(C.op, c, e, w, n, s, nx, ny, j, i, (T)C.dx, (T)C.dy);
  const size_t k = (size_t)j * nx + i;
  ((T*)C.w0)[k] = wv;
  const T rr = R[k];
  acc[0] += (double)rr * (double)c;
  acc[1] += (double)wv * (double)c;
  acc[2] += (double)rr * (double)rr;
}

template <class T>
__device__ inline void cell_cg(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const int nx = C.nx, ny = C.ny;
  const int cur = A.it & 1;
  const T* R = (const T*)(cur ? C.r1 : C.r);
  const T* W = (const T*)(cur ? C.w1 : C.w0);
  const T* Sold = (const T*)(cur ? C.d0 : C.d1);
  T* Snew = (T*)(cur ? C.d1 : C.d0);
  T* Rn = (T*)(cur ? C.r : C.r1);
  T* Wn = (T*)(cur ? C.w0 : C.w1);
  T* P = (T*)C.pd;
  T* X = (T*)C.x;
  const int first = (A.it == 0);
  const T al = (T)A.alpha, be = (T)A.beta;
  auto minv = [&](int jj, int ii) -> T {
    if constexpr (sizeof(T) == sizeof(float)) return minv_f32(C, A, jj, ii);
    else return (T)jacobi_minv(C.jacobi, C.invd, nx, ny, jj, ii);
  };
  // s_i and r_{i+1} at (jj, ii)
  auto snew = [&](size_t kk) -> T { return first ? W[kk] : (T)(W[kk] + be * Sold[kk]); };
  auto unext = [&](int jj, int ii, T* s_out) -> T {
    const size_t kk = (size_t)jj * nx + ii;
    const T sv = snew(kk);
    if (s_out) *s_out = sv;
    const T rv = R[kk] - al * sv;
    return minv(jj, ii) * rv;
  };
  const size_t k = (size_t)j * nx + i;
  T s_c;
  const T c = unext(j, i, &s_c);
  const T e = (i + 1 < nx) ? unext(j, i + 1, nullptr) : T(0);
  const T w = (i > 0) ? unext(j, i - 1, nullptr) : T(0);
  const T n = (j + 1 < ny) ? unext(j + 1, i, nullptr) : T(0);
  const T s = (j > 0) ? unext(j - 1, i, nullptr) : T(0);
  T wv;
  if constexpr (sizeof(T) == sizeof(float)) wv = apply_op5_fast(C.op, c, e, w, n, s, nx, ny, j, i, A.cx, A.cy);
  else wv = apply_op5<T>(C.op, c, e, w, n, s, nx, ny, j, i, (T)C.dx, (T)C.dy);
  const T r_c = R[k] - al * s_c;
  const T u_i = minv(j, i) * R[k];
  const T p_i = first ? u_i : (T)(u_i + be * P[k]);
  X[k] = X[k] + al * p_i;
  P[k] = p_i;
  Snew[k] = s_c;
  Rn[k] = r_c;
  Wn[k] = wv;
  acc[0] += (double)r_c * (double)c;
  acc[1] += (double)wv * (double)c;
  acc[2] += (double)r_c * (double)r_c;
}

// Float32 CG iteration over a workgroup's whole region (64 columns x 16*M
// rows) in one pass. Every thread computes u_{i+1} = M r_{i+1} at its M cells
// (rows 16*m + threadIdx.y) and threads 0..2*16*M+127 also at the 1-cell halo
// (two rows, two columns) into LDS; after one barrier each thread applies the
// stencil from LDS and stores s_i, r_{i+1}, w_{i+1}, p_i, x_{i+1}. All global
// loads of a thread are independent and issued together.
constexpr int LW = TX + 2;  // LDS row pitch: 64 columns + 1-cell halo each side
template <int M>
__device__ inline void region_cg_f32(const PfCav& C, const PhaseArgs& A, int b, double* acc, float* __restrict__ lu) {
  const int nx = C.nx, ny = C.ny;
  const int cur = A.it & 1;
  const float* __restrict__ R = (const float*)(cur ? C.r1 : C.r);
  const float* __restrict__ W = (const float*)(cur ? C.w1 : C.w0);
  const float* __restrict__ Sold = (const float*)(cur ? C.d0 : C.d1);
  float* __restrict__ Snew = (float*)(cur ? C.d1 : C.d0);
  float* __restrict__ Rn = (float*)(cur ? C.r : C.r1);
  float* __restrict__ Wn = (float*)(cur ? C.w0 : C.w1);
  float* __restrict__ P = (float*)C.pd;
  float* __restrict__ X = (float*)C.x;
  const int first = (A.it == 0);
  const float al = A.alphaf, be = A.betaf;
  const int li = threadIdx.x, lj = threadIdx.y;
  const int tid = lj * TX + li;
  const int i0 = (b % C.tiles_x) * TX;
  const int j0 = (b / C.tiles_x) * (TY * M);
  const int i = i0 + li;
  constexpr int RH = TY * M;
  // Pass 1: issue every global load of this thread (M cells + at most one
  // halo cell) before any store, so they are all in flight together.
  float wv_[M], so_[M], ro_[M], po_[M], xo_[M];
  bool in[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    in[m] = (i < nx) && (j < ny);
    wv_[m] = so_[m] = ro_[m] = po_[m] = xo_[m] = 0.f;
    if (in[m]) {
      const size_t k = (size_t)j * nx + i;
      wv_[m] = W[k];
      ro_[m] = R[k];
      xo_[m] = X[k];
      if (!first) {
        so_[m] = Sold[k];
        po_[m] = P[k];
      }
    }
  }
  // halo: rows j0-1 and j0+RH (threads 0..127), columns i0-1 and i0+64
  // (threads 128..128+2*RH-1)
  int hslot = -1, hj = 0, hi = 0;
  if (tid < 2 * TX) {
    const int side = tid / TX;
    hi = i0 + (tid % TX);
    hj = side ? j0 + RH : j0 - 1;
    hslot = (side ? RH + 1 : 0) * LW + (tid % TX) + 1;
  } else if (tid < 2 * TX + 2 * RH) {
    const int q = tid - 2 * TX;
    const int side = q / RH, row = q % RH;
    hj = j0 + row;
    hi = side ? i0 + TX : i0 - 1;
    hslot = (row + 1) * LW + (side ? TX + 1 : 0);
  }
  const bool hin = hslot >= 0 && hi >= 0 && hi < nx && hj >= 0 && hj < ny;
  float hw = 0.f, hs = 0.f, hr = 0.f;
  if (hin) {
    const size_t kk = (size_t)hj * nx + hi;
    hw = W[kk];
    hr = R[kk];
    if (!first) hs = Sold[kk];
  }
  // Pass 2: this thread's cells, then the halo value, into LDS.
  float rc[M], uc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    rc[m] = uc[m] = 0.f;
    const int j = j0 + TY * m + lj;
    if (in[m]) {
      const size_t k = (size_t)j * nx + i;
      const float mi = minv_f32(C, A, j, i);
      const float sv = first ? wv_[m] : (wv_[m] + be * so_[m]);
      rc[m] = ro_[m] - al * sv;
      uc[m] = mi * rc[m];
      const float ui = mi * ro_[m];
      const float p_i = first ? ui : (ui + be * po_[m]);
      X[k] = xo_[m] + al * p_i;
      P[k] = p_i;
      Snew[k] = sv;
      Rn[k] = rc[m];
    }
    lu[(TY * m + lj + 1) * LW + li + 1] = uc[m];
  }
  if (hslot >= 0) {
    float hu = 0.f;
    if (hin) {
      const float sv = first ? hw : (hw + be * hs);
      hu = minv_f32(C, A, hj, hi) * (hr - al * sv);
    }
    lu[hslot] = hu;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (in[m]) {
      const int j = j0 + TY * m + lj;
      const int rr = TY * m + lj + 1;
      const float c = uc[m];
      const float e = lu[rr * LW + li + 2];
      const float w = lu[rr * LW + li];
      const float n = lu[(rr + 1) * LW + li + 1];
      const float s = lu[(rr - 1) * LW + li + 1];
      const float wv = apply_op5_fast(C.op, c, e, w, n, s, nx, ny, j, i, A.cx, A.cy);
      Wn[(size_t)j * nx + i] = wv;
      acc[0] += (double)rc[m] * (double)c;
      acc[1] += (double)wv * (double)c;
      acc[2] += (double)rc[m] * (double)rc[m];
    }
  }
}

__device__ inline void region_cg_f32_dispatch(const PfCav& C, const PhaseArgs& A, int b, double* acc, float* lu) {
  switch (C.tpb) {
    case 1: region_cg_f32<1>(C, A, b, acc, lu); break;
    case 2: region_cg_f32<2>(C, A, b, acc, lu); break;
    case 4: region_cg_f32<4>(C, A, b, acc, lu); break;
    default: region_cg_f32<8>(C, A, b, acc, lu); break;
  }
}

template <class T>
__device__ inline void cell_phase(int phase, const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  switch (phase) {
    case PH_STATS: cell_stats<T>(C, j, i, acc); break;
    case PH_PRED_A: cell_pred_a<T>(C, A, j, i); break;
    case PH_PRED_B: cell_pred_b<T>(C, A, j, i, acc); break;
    case PH_PINIT: cell_pinit<T>(C, A, j, i, acc); break;
    case PH_P: cell_p<T>(C, A, j, i, acc); break;
    case PH_Q: cell_q<T>(C, A, j, i, acc); break;
    case PH_OUTER_A: cell_outer_a<T>(C, j, i); break;
    case PH_OUTER_B: cell_outer_b<T>(C, j, i, acc); break;
    case PH_FIN: cell_fin<T>(C, A, j, i, acc); break;
    case PH_CORR_A: cell_corr_a<T>(C, A, j, i, acc); break;
    case PH_CORR_B: cell_corr_b<T>(C, j, i, acc); break;
    case PH_CG0: cell_cg0<T>(C, A, j, i, acc); break;
    case PH_CG: cell_cg<T>(C, A, j, i, acc); break;
    default: break;
  }
}

// One round over every workgroup of every cavity. `parity` = round & 1.
// U = tiles per unrolled chunk in the P and Q loops.
// Kernel arguments. A batch holding a single cavity passes that cavity's
// descriptor by value (SINGLE), so the first memory round trip of every
// workgroup already fetches the state and the partials; multi-cavity batches
// look the cavity up through blk_cav.
struct KArgs {
  const PfCav* cavs;
  PfScal* scal;
  const int* blk_cav;
  int* active;
  int max_cav;
  int parity;
  int slot0;   // SINGLE: the cavity's slot
  int pad;
  PfCav cav0;  // SINGLE: the cavity's descriptor
};

template <class T, int U, bool SINGLE>
__global__ __launch_bounds__(NTH) void k_round(const KArgs KA) {
  __shared__ PfScal st;
  __shared__ double sm[NWAVE][NP];
  __shared__ double tot[NP];
  extern __shared__ float lds_dyn[];  // (16 * tpb + 2) * LW floats (float32 CG region)
  const int g = blockIdx.x;
  int c;
  const PfCav* cp;
  if constexpr (SINGLE) {
    c = KA.slot0;
    cp = &KA.cav0;
  } else {
    c = KA.blk_cav[g];
    if (c < 0) return;
    cp = KA.cavs + c;
  }
  const PfCav& C = *cp;
  const int parity = KA.parity;
  const int max_cav = KA.max_cav;
  PfScal* scal = KA.scal;
  int* active = KA.active;
  // XCD-aware region order: workgroups are dispatched round-robin over the 8
  // XCDs (g mod 8), each with its own L2. When the cavity's workgroup range is
  // 8-aligned, XCD x gets the contiguous regions [x*nblk/8, (x+1)*nblk/8), so
  // the halo rows/columns a region reads were mostly fetched into the same L2
  // by its neighbours. Partials are indexed by region, so the reduction order
  // does not depend on the mapping.
  const int lg = g - C.blk0;
  const int b = ((C.blk0 & 7) == 0 && (C.nblk & 7) == 0) ? (lg & 7) * (C.nblk >> 3) + (lg >> 3) : lg;
  const int tid = tid_flat();
  const uint32_t* src = (const uint32_t*)(scal + (size_t)(parity ^ 1) * max_cav + c);
  uint32_t* dst = (uint32_t*)(scal + (size_t)parity * max_cav + c);
  constexpr int NW32 = (int)(sizeof(PfScal) / 4);
  // One memory round trip: the state words and every slot of the previous
  // round's partials are loaded together (which slots matter is only known
  // once the state has arrived).
  const double* prev = C.part + (size_t)(parity ^ 1) * C.nblk * NP;
  double pa[NP] = {0.0, 0.0, 0.0, 0.0};
  const int nblk = C.nblk;
  double pq0[NP];
  const bool have0 = tid < nblk;
#pragma unroll
  for (int k = 0; k < NP; ++k) pq0[k] = have0 ? prev[(size_t)tid * NP + k] : 0.0;
  for (int w = tid; w < NW32; w += NTH) ((uint32_t*)&st)[w] = src[w];
  __syncthreads();
  if (st.pending != PH_IDLE) {
    const int K = phase_nparts(st.pending);
    const unsigned mm = phase_maxmask(st.pending);
    if (K > 0) {
      if (have0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) pa[k] = pq0[k];
      }
      for (int q = tid + NTH; q < nblk; q += NTH) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
          if (k < K) pa[k] = redop(mm, k, pa[k], prev[(size_t)q * NP + k]);
      }
      block_reduce_phase(st.pending, pa, sm, tot);
    }
    if (tid == 0) tail(C, st, tot);
    __syncthreads();  // tail result visible; sm and tot free for reuse
  }
  const int phase = st.phase;
  if (phase <= PH_IDLE || phase >= PH_DONE) {
    if (b == 0) {
      const int newly_done = st.done_now;
      __syncthreads();
      if (tid == 0) st.done_now = 0;
      __syncthreads();
      for (int w = tid; w < NW32; w += NTH) dst[w] = ((const uint32_t*)&st)[w];
      if (tid == 0 && newly_done) atomicSub(active, 1);
    }
    return;
  }
  PhaseArgs A;
  A.dt = st.dt;
  A.mean = st.mean;
  A.beta = st.beta;
  A.alpha = st.alpha;
  A.it = st.it;
  A.x_is_b = st.x_is_b;
  A.x0 = 0.0;
  if (phase == PH_FIN) A.x0 = st.x_is_b ? C.b[0] : C.x64[0];
  A.cx = (float)(1.0 / (C.dx * C.dx));
  A.cy = (float)(1.0 / (C.dy * C.dy));
  A.betaf = (float)A.beta;
  A.alphaf = (float)A.alpha;
  A.minvf = (float)C.invd;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0};
  const int ty0 = (b / C.tiles_x) * C.tpb;
  TileIt ti{b % C.tiles_x, ty0, min(ty0 + C.tpb, C.tiles_y)};
  if (phase == PH_P) {
    for (int m = 0; m < C.tpb; m += U) chunk_p<T, U>(C, A, ti, acc);
  } else if (phase == PH_Q) {
    for (int m = 0; m < C.tpb; m += U) chunk_q<T, U>(C, A, ti, acc);
  } else if (sizeof(T) == sizeof(float) && phase == PH_CG) {
    region_cg_f32_dispatch(C, A, b, acc, lds_dyn);
  } else {
    for (int ty = ti.ty; ty < ti.tyend; ++ty) {
      const int i = ti.tx * TX + threadIdx.x;
      const int j = ty * TY + threadIdx.y;
      if (i < C.nx && j < C.ny) cell_phase<T>(phase, C, A, j, i, acc);
    }
  }
  const int K = phase_nparts(phase);
  if (K > 0) {
    block_reduce_phase(phase, acc, sm, tot);
    if (tid < K) C.part[((size_t)parity * C.nblk + b) * NP + tid] = tot[tid];
  }
  if (b == 0) {
    __syncthreads();
    if (tid == 0) st.pending = phase;
    __syncthreads();
    for (int w = tid; w < NW32; w += NTH) dst[w] = ((const uint32_t*)&st)[w];
  }
}

// Standalone float64 operators (numerics API), one cell per thread.
__global__ void k_op(int opk, int nx, int ny, double dx, double dy, const double* __restrict__ in0,
                     const double* __restrict__ in1, const double* __restrict__ in2, double* __restrict__ out0,
                     double* __restrict__ out1) {
  const int i = blockIdx.x * TX + threadIdx.x;
  const int j = blockIdx.y * TY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const size_t k = (size_t)j * nx + i;
  Grid<double> A{in0, nx}, B{in1, nx}, Cc{in2, nx};
  switch (opk) {
    case PF_OPK_GRAD_COHERENT:
      out0[k] = grad_x_coherent<double>(A, nx, ny, j, i, dx);
      out1[k] = grad_y_coherent<double>(A, nx, ny, j, i, dy);
      break;
    case PF_OPK_DIV_COHERENT: out0[k] = div_coherent<double>(A, B, nx, ny, j, i, dx, dy); break;
    case PF_OPK_LAP_COHERENT: {
      const double c = A(j, i);
      const double e = (i + 1 < nx) ? A(j, i + 1) : 0.0;
      const double w = (i > 0) ? A(j, i - 1) : 0.0;
      const double n = (j + 1 < ny) ? A(j + 1, i) : 0.0;
      const double s = (j > 0) ? A(j - 1, i) : 0.0;
      out0[k] = lap_coherent5<double>(c, e, w, n, s, nx, ny, j, i, dx, dy);
      break;
    }
    case PF_OPK_DIV_CENTRAL: out0[k] = div_central<double>(A, B, nx, ny, j, i, dx, dy); break;
    case PF_OPK_ADV_QUICK: out0[k] = adv_quick<double>(A(j, i), B(j, i), Cc, nx, ny, j, i, dx, dy); break;
    case PF_OPK_ADV_UPWIND: out0[k] = adv_upwind<double>(A(j, i), B(j, i), Cc, nx, ny, j, i, dx, dy); break;
    case PF_OPK_NEG_LAP_PINNED: out0[k] = apply_op_at<double>(0, A, nx, ny, j, i, dx, dy); break;
    case PF_OPK_POS_LAP_DIRICHLET: out0[k] = apply_op_at<double>(1, A, nx, ny, j, i, dx, dy); break;
    default: break;
  }
}

template <class T>
__global__ void k_f64_to(const double* __restrict__ src, T* __restrict__ dst, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = (T)src[k];
}
template <class T>
__global__ void k_to_f64(const T* __restrict__ src, double* __restrict__ dst, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = (double)src[k];
}

}  // namespace pf
