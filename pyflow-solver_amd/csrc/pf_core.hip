// libpyflow_hip: gfx950 HIP implementation of PyFlow's projection time step.
//
// Execution model
// ---------------
// A batch holds cavities (independent lid-driven-cavity problems of any size).
// Each cavity is cut into 64 x TY-cell tiles; one workgroup (64 x 4 threads,
// one wave per row of 64 cells) owns one tile. Every cavity runs a phase
// machine on the device:
//
//   STATS -> PRED_A -> PRED_B -> PINIT -> {P -> Q}* [-> OUTER_A -> OUTER_B]*
//         -> FIN -> CORR_A -> CORR_B -> (STATS of the next step | DONE)
//
// A "round" is one launch over all tiles of all cavities; each tile executes
// the phase its cavity is in. Phases only read data written in earlier rounds
// (plus their own cells), so no tile waits on another inside a launch: there
// is no spin-wait anywhere, and a round always drains. Per-cavity reductions
// are deterministic: each tile writes fixed-order FP64 partials, the last tile
// to arrive (atomic ticket) sums them in tile order and runs the cavity's
// scalar tail (dt controller, CG alpha/beta, convergence tests), then advances
// the cavity's phase for the next round.
//
// Reference correspondence (beautifulplanet/PyFlow-Solver, src/pyflow):
//   STATS   solvers/solver.py:25-62      abs-max, central divergence, dt control
//   PRED_A  solvers/solver.py:73-95      u* = u - dt*conv (QUICK / upwind / none)
//   PRED_B  solvers/pressure_solver.py:116-127  rhs = -div_c(u*)/dt, mean
//   PINIT   pressure_solver.py:126-128; scipy cg setup (x0=0, r=b, atol)
//   P, Q    scipy cg loop body (iterative.py:392-415): p = z + beta p; q = A p;
//           alpha = rho/(p.q); x += alpha p; r -= alpha q
//   OUTER_* (MIXED only) FP64 iterative refinement: x64 += e; r = b - A x64
//   FIN     linear_solvers/interface.py:47-52 true residual; pressure_solver.py:147-149 p
//   CORR_A  pressure_solver.py:153-157 correction; solver.py:103-112 lid BC, Ru/Rv, NaN
//   CORR_B  pressure_solver.py:165-171 div_c after; solver.py:114 continuity
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pf_cells.hpp"
#include "pyflow_hip.h"

namespace pf {

constexpr int TX = 64;        // tile width = wavefront width
constexpr int TY = 4;         // tile height (rows); one row per wave
constexpr int NTH = TX * TY;  // threads per workgroup
constexpr int NWAVE = NTH / 64;
constexpr int NP = 4;         // FP64 partials per tile

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
  PH_DONE = 12,
  PH_COUNT = 13
};

// Read-mostly description of one cavity (rewritten only between runs).
struct PfCav {
  int nx, ny, tiles_x, ntiles, tile0;
  int scheme, lid_on, jacobi, maxiter, op, mode, prec;
  double lid, dx, dy, nu, cfl, growth, tol, invd, inner_rtol;
  void *u, *v, *p, *us, *vs, *r, *d0, *d1, *x;  // element type T
  double *b, *x64;                              // x64 aliases x in FP64 mode
  double *part;                                 // ntiles * NP partials
};

// Dynamic per-cavity scalars, owned by the device during a run.
struct PfScal {
  int phase, steps_left, steps_done, it, it_total, info, outer, x_is_b;
  int hist_len, safeguard, nan, converged;
  unsigned ticket;
  int pad[3];
  double dt_prev, dt, cfl_now, umax, vmax, hist[5], div_before;
  double mean, div_initial, bnorm, atol, inner_atol, rho, rho_prev, beta, pq, alpha, rnorm2, rn_outer;
  double rnorm_true, div_after, ru, rv, continuity;
};

template <class T>
struct Grid {
  const T* __restrict__ a;
  int nx;
  __device__ T operator()(int j, int i) const { return a[(size_t)j * nx + i]; }
};

__device__ inline int tid_flat() { return threadIdx.y * TX + threadIdx.x; }

// Reduce K per-thread values over the workgroup (sum, or NaN-propagating max
// where bit k of maxmask is set). Fixed butterfly + fixed wave order, so the
// result is bitwise reproducible. Result valid in thread 0.
template <int K>
__device__ inline void block_reduce(double (&v)[K], unsigned maxmask, double (*sm)[NP]) {
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double o = __shfl_xor(v[k], off, 64);
      v[k] = ((maxmask >> k) & 1u) ? nanmax(v[k], o) : v[k] + o;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[wave][k] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = sm[0][k];
      for (int w = 1; w < NWAVE; ++w) a = ((maxmask >> k) & 1u) ? nanmax(a, sm[w][k]) : a + sm[w][k];
      v[k] = a;
    }
  }
  __syncthreads();
}

// Write this tile's partials, take a ticket; the last tile of the cavity sums
// all partials in tile order into tot[] (valid in thread 0) and returns true.
template <int K>
__device__ bool reduce_arrive(const PfCav& C, PfScal* S, int lt, double (&v)[K], unsigned maxmask,
                              double (&tot)[K]) {
  __shared__ double sm[NWAVE][NP];
  __shared__ int is_last;
  const int tid = tid_flat();
  block_reduce<K>(v, maxmask, sm);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) C.part[(size_t)lt * NP + k] = v[k];
    __threadfence();
    const unsigned tk = atomicAdd(&S->ticket, 1u);
    is_last = (tk == (unsigned)(C.ntiles - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  for (int t = tid; t < C.ntiles; t += NTH) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double q = C.part[(size_t)t * NP + k];
      a[k] = ((maxmask >> k) & 1u) ? nanmax(a[k], q) : a[k] + q;
    }
  }
  block_reduce<K>(a, maxmask, sm);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] = a[k];
    S->ticket = 0u;
  }
  return true;
}

// Arrival without partials: returns true in every thread of the last tile.
__device__ bool arrive_only(const PfCav& C, PfScal* S) {
  __shared__ int is_last;
  const int tid = tid_flat();
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const unsigned tk = atomicAdd(&S->ticket, 1u);
    is_last = (tk == (unsigned)(C.ntiles - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  if (tid == 0) S->ticket = 0u;
  return true;
}

__device__ inline void enter_done(PfScal* S, int* active) {
  S->phase = PH_DONE;
  atomicSub(active, 1);
}

// ---------------------------------------------------------------------------
// Phase bodies. Each is executed by every thread of a tile; (j, i) is this
// thread's cell, `in` whether it lies inside the cavity.
// ---------------------------------------------------------------------------
template <class T>
__device__ void ph_stats(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
  double v[3] = {0.0, 0.0, 0.0};
  if (in) {
    v[0] = fabs((double)U(j, i));
    v[1] = fabs((double)V(j, i));
    const double cd = (double)div_central<T>(U, V, C.nx, C.ny, j, i, T(1), T(1));
    v[2] = cd * cd;
  }
  double tot[3];
  if (!reduce_arrive<3>(C, S, lt, v, 0x3u, tot)) return;
  if (tid_flat() != 0) return;
  // dt controller, solver.py:25-62 (Python float semantics)
  const double umax = tot[0] + 1e-12;
  const double vmax = tot[1] + 1e-12;
  const double dt_cfl = pymin(C.cfl * C.dx / umax, C.cfl * C.dy / vmax);
  const double dt_diff = 0.25 * pymin(C.dx * C.dx, C.dy * C.dy) / pymax(C.nu, 1e-12);
  const double dt_cand = pymin(dt_cfl, dt_diff);
  const double cap = S->dt_prev * C.growth;
  double dt = (dt_cand > cap) ? cap : dt_cand;
  const double div0 = sqrt(tot[2]);
  int hl = S->hist_len;
  if (hl == 5) {
    for (int k = 0; k < 4; ++k) S->hist[k] = S->hist[k + 1];
    S->hist[4] = div0;
  } else {
    S->hist[hl] = div0;
    hl += 1;
    S->hist_len = hl;
  }
  int sg = 0;
  if (hl >= 3) {
    const double a = S->hist[hl - 3], b = S->hist[hl - 2], c = S->hist[hl - 1];
    if (a > 0 && b > a * 1.01 && c > b * 1.01) {
      dt *= 0.5;
      sg = 1;
    }
  }
  S->safeguard = sg;
  S->dt = dt;
  S->dt_prev = dt;
  S->umax = umax;
  S->vmax = vmax;
  S->div_before = div0;
  S->cfl_now = pymax(umax * dt / C.dx, vmax * dt / C.dy);
  S->phase = PH_PRED_A;
}

template <class T>
__device__ void ph_pred_a(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  if (in) {
    Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
    const size_t k = (size_t)j * C.nx + i;
    const T uc = U(j, i), vc = V(j, i);
    T us = uc, vs = vc;
    if (C.scheme != PF_SCHEME_NONE) {
      const T dt = (T)S->dt, dx = (T)C.dx, dy = (T)C.dy;
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
  (void)lt;
  if (!arrive_only(C, S)) return;
  if (tid_flat() == 0) S->phase = PH_PRED_B;
}

template <class T>
__device__ void ph_pred_b(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  double v[2] = {0.0, 0.0};
  if (in) {
    Grid<T> US{(const T*)C.us, C.nx}, VS{(const T*)C.vs, C.nx};
    const T dv = div_coherent<T>(US, VS, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
    const double rhs = (-(double)dv) / pymax(S->dt, 1e-14);
    C.b[(size_t)j * C.nx + i] = rhs;
    v[0] = rhs;
    v[1] = (double)dv * (double)dv;
  }
  double tot[2];
  if (!reduce_arrive<2>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  S->mean = tot[0] / (double)((size_t)C.nx * C.ny);
  S->div_initial = sqrt(tot[1]);
  S->phase = PH_PINIT;
}

// Decide what follows a completed inner iteration / initialisation.
__device__ inline void cg_next(const PfCav& C, PfScal* S, int* active_unused) {
  (void)active_unused;
  const bool mixed = (C.prec == PF_PREC_MIXED);
  if (S->it_total >= C.maxiter) {
    S->info = C.maxiter;
    S->phase = mixed ? PH_OUTER_A : PH_FIN;
    return;
  }
  if (sqrt(S->rnorm2) < S->inner_atol) {
    if (!mixed) S->info = 0;
    S->phase = mixed ? PH_OUTER_A : PH_FIN;
    return;
  }
  S->phase = PH_P;
}

template <class T>
__device__ void ph_pinit(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  double v[3] = {0.0, 0.0, 0.0};
  if (in) {
    const size_t k = (size_t)j * C.nx + i;
    double bb = C.b[k];
    if (C.mode != PF_MODE_SOLVE) {
      bb = bb - S->mean;
      if (k == 0) bb = 0.0;
      C.b[k] = bb;
    }
    const T rr = (T)bb;
    ((T*)C.r)[k] = rr;
    ((T*)C.x)[k] = T(0);
    if (C.prec == PF_PREC_MIXED) C.x64[k] = 0.0;
    const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rr;
    v[0] = bb * bb;
    v[1] = (double)rr * (double)z;
    v[2] = (double)rr * (double)rr;
  }
  double tot[3];
  if (!reduce_arrive<3>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  const double bn = sqrt(tot[0]);
  S->bnorm = bn;
  S->atol = pymax(0.0, C.tol * bn);
  S->it = 0;
  S->it_total = 0;
  S->outer = 0;
  S->info = 0;
  S->x_is_b = 0;
  if (bn == 0.0) {  // scipy returns x = b, info 0
    S->x_is_b = 1;
    S->phase = PH_FIN;
    return;
  }
  S->rho = tot[1];
  S->rnorm2 = tot[2];
  S->rn_outer = bn;
  if (C.maxiter <= 0) {  // loop never runs: info = maxiter (= 0 or negative)
    S->info = C.maxiter;
    S->phase = PH_FIN;
    return;
  }
  if (C.prec == PF_PREC_MIXED) {
    if (bn < S->atol) {
      S->phase = PH_FIN;
      return;
    }
    S->inner_atol = pymax(C.inner_rtol * bn, 0.5 * S->atol);
    S->phase = PH_P;
    return;
  }
  S->inner_atol = S->atol;
  if (sqrt(S->rnorm2) < S->atol) {
    S->phase = PH_FIN;
    return;
  }
  S->phase = PH_P;
}

// Search direction: d = z (first inner iteration) or (d_old*beta) + z.
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
__device__ void ph_p(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  const int it = S->it;
  const T* R = (const T*)C.r;
  const T* Dold = (const T*)((it & 1) ? C.d0 : C.d1);
  T* Dnew = (T*)((it & 1) ? C.d1 : C.d0);
  const int first = (it == 0);
  const double beta = S->beta;
  double v[1] = {0.0};
  if (in) {
    const int nx = C.nx, ny = C.ny;
    const T c = dnew_at<T>(C, R, Dold, first, beta, j, i);
    const T e = (i + 1 < nx) ? dnew_at<T>(C, R, Dold, first, beta, j, i + 1) : T(0);
    const T w = (i > 0) ? dnew_at<T>(C, R, Dold, first, beta, j, i - 1) : T(0);
    const T n = (j + 1 < ny) ? dnew_at<T>(C, R, Dold, first, beta, j + 1, i) : T(0);
    const T s = (j > 0) ? dnew_at<T>(C, R, Dold, first, beta, j - 1, i) : T(0);
    Dnew[(size_t)j * nx + i] = c;
    const T q = apply_op5<T>(C.op, c, e, w, n, s, nx, ny, j, i, (T)C.dx, (T)C.dy);
    v[0] = (double)c * (double)q;
  }
  double tot[1];
  if (!reduce_arrive<1>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  S->pq = tot[0];
  S->alpha = S->rho / tot[0];
  S->phase = PH_Q;
}

template <class T>
__device__ void ph_q(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  const int it = S->it;
  const T* D = (const T*)((it & 1) ? C.d1 : C.d0);
  const double alpha = S->alpha;
  double v[2] = {0.0, 0.0};
  if (in) {
    Grid<T> DG{D, C.nx};
    const size_t k = (size_t)j * C.nx + i;
    const T q = apply_op_at<T>(C.op, DG, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
    T* X = (T*)C.x;
    T* R = (T*)C.r;
    const T d = D[k];
    X[k] = (T)((double)X[k] + alpha * (double)d);
    const T rn = (T)((double)R[k] - alpha * (double)q);
    R[k] = rn;
    const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rn;
    v[0] = (double)rn * (double)z;
    v[1] = (double)rn * (double)rn;
  }
  double tot[2];
  if (!reduce_arrive<2>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  S->rho_prev = S->rho;
  S->rho = tot[0];
  S->rnorm2 = tot[1];
  S->beta = S->rho / S->rho_prev;
  S->it += 1;
  S->it_total += 1;
  cg_next(C, S, nullptr);
}

// MIXED: fold the inner correction into the FP64 solution.
template <class T>
__device__ void ph_outer_a(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  if (in) {
    const size_t k = (size_t)j * C.nx + i;
    T* X = (T*)C.x;
    C.x64[k] = C.x64[k] + (double)X[k];
    X[k] = T(0);
  }
  (void)lt;
  if (!arrive_only(C, S)) return;
  if (tid_flat() == 0) S->phase = PH_OUTER_B;
}

// MIXED: FP64 residual of the refined solution, restart the FP32 inner PCG.
template <class T>
__device__ void ph_outer_b(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  double v[3] = {0.0, 0.0, 0.0};
  if (in) {
    const size_t k = (size_t)j * C.nx + i;
    Grid<double> XG{C.x64, C.nx};
    const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
    const double r64 = C.b[k] - ax;
    const T rt = (T)r64;
    ((T*)C.r)[k] = rt;
    const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rt;
    v[0] = r64 * r64;
    v[1] = (double)rt * (double)z;
    v[2] = (double)rt * (double)rt;
  }
  double tot[3];
  if (!reduce_arrive<3>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  const double rn = sqrt(tot[0]);
  S->rn_outer = rn;
  S->outer += 1;
  if (S->info != 0) {  // iteration budget exhausted
    S->phase = PH_FIN;
    return;
  }
  if (rn < S->atol) {
    S->info = 0;
    S->phase = PH_FIN;
    return;
  }
  S->rho = tot[1];
  S->rnorm2 = tot[2];
  S->it = 0;
  S->inner_atol = pymax(C.inner_rtol * rn, 0.5 * S->atol);
  S->phase = PH_P;
}

template <class T>
__device__ void ph_fin(const PfCav& C, PfScal* S, int lt, int j, int i, bool in, int* active) {
  double v[1] = {0.0};
  if (in) {
    const size_t k = (size_t)j * C.nx + i;
    const double* XS = S->x_is_b ? C.b : C.x64;
    Grid<double> XG{XS, C.nx};
    const double xv = XS[k];
    const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
    const double rr = C.b[k] - ax;
    v[0] = rr * rr;
    const double pv = (C.mode == PF_MODE_SOLVE) ? xv : (xv - XS[0]);
    ((T*)C.p)[k] = (T)pv;
  }
  double tot[1];
  if (!reduce_arrive<1>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  const double rn = sqrt(tot[0]);
  S->rnorm_true = rn;
  const double bn = (S->bnorm == 0.0) ? 1.0 : S->bnorm;
  S->converged = (S->info == 0) && (rn <= C.tol * bn);
  if (C.mode == PF_MODE_SOLVE) {
    S->steps_done += 1;
    enter_done(S, active);
  } else {
    S->phase = PH_CORR_A;
  }
}

template <class T>
__device__ void ph_corr_a(const PfCav& C, PfScal* S, int lt, int j, int i, bool in) {
  double v[3] = {0.0, 0.0, 0.0};
  if (in) {
    const int nx = C.nx, ny = C.ny;
    const size_t k = (size_t)j * nx + i;
    Grid<T> PG{(const T*)C.p, nx};
    const T dt = (T)S->dt;
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
    v[0] = du * du;
    v[1] = dv * dv;
    const bool bad = !isfinite((double)uf) || !isfinite((double)vc) || !isfinite((double)PG(j, i));
    v[2] = bad ? 1.0 : 0.0;
  }
  double tot[3];
  if (!reduce_arrive<3>(C, S, lt, v, 0x4u, tot)) return;
  if (tid_flat() != 0) return;
  const double nn = (double)((size_t)C.nx * C.ny);
  const double den = pymax(1.0, nn);
  S->ru = sqrt(tot[0]) / den;
  S->rv = sqrt(tot[1]) / den;
  S->nan = (tot[2] != 0.0) ? 1 : 0;
  S->phase = PH_CORR_B;
}

template <class T>
__device__ void ph_corr_b(const PfCav& C, PfScal* S, int lt, int j, int i, bool in, int* active) {
  double v[2] = {0.0, 0.0};
  if (in) {
    Grid<T> US{(const T*)C.us, C.nx}, VS{(const T*)C.vs, C.nx};
    Grid<T> U{(const T*)C.u, C.nx}, V{(const T*)C.v, C.nx};
    const double dc = (double)div_coherent<T>(US, VS, C.nx, C.ny, j, i, (T)C.dx, (T)C.dy);
    const double cd = (double)div_central<T>(U, V, C.nx, C.ny, j, i, T(1), T(1));
    v[0] = dc * dc;
    v[1] = cd * cd;
  }
  double tot[2];
  if (!reduce_arrive<2>(C, S, lt, v, 0u, tot)) return;
  if (tid_flat() != 0) return;
  S->div_after = sqrt(tot[0]);
  S->continuity = sqrt(tot[1]) / pymax(1.0, (double)((size_t)C.nx * C.ny));
  S->steps_done += 1;
  S->steps_left -= 1;
  if (C.mode == PF_MODE_STEP && S->steps_left > 0) {
    S->phase = PH_STATS;
  } else {
    enter_done(S, active);
  }
}

// One round: every tile executes its cavity's current phase. PH selects a
// single phase at compile time (0 = dispatch on the phase read at entry).
template <class T, int PH>
__global__ __launch_bounds__(NTH) void k_round(const PfCav* __restrict__ cavs, PfScal* __restrict__ scal,
                                               const int* __restrict__ tile_cav, int* __restrict__ active) {
  const int t = blockIdx.x;
  const int c = tile_cav[t];
  if (c < 0) return;
  const PfCav& C = cavs[c];
  PfScal* S = scal + c;
  const int phase = S->phase;
  if (PH != 0 && phase != PH) return;
  if (phase <= PH_IDLE || phase >= PH_DONE) return;
  const int lt = t - C.tile0;
  const int tx = lt % C.tiles_x, ty = lt / C.tiles_x;
  const int i = tx * TX + threadIdx.x;
  const int j = ty * TY + threadIdx.y;
  const bool in = (i < C.nx) && (j < C.ny);
  switch (phase) {
    case PH_STATS: ph_stats<T>(C, S, lt, j, i, in); break;
    case PH_PRED_A: ph_pred_a<T>(C, S, lt, j, i, in); break;
    case PH_PRED_B: ph_pred_b<T>(C, S, lt, j, i, in); break;
    case PH_PINIT: ph_pinit<T>(C, S, lt, j, i, in); break;
    case PH_P: ph_p<T>(C, S, lt, j, i, in); break;
    case PH_Q: ph_q<T>(C, S, lt, j, i, in); break;
    case PH_OUTER_A: ph_outer_a<T>(C, S, lt, j, i, in); break;
    case PH_OUTER_B: ph_outer_b<T>(C, S, lt, j, i, in); break;
    case PH_FIN: ph_fin<T>(C, S, lt, j, i, in, active); break;
    case PH_CORR_A: ph_corr_a<T>(C, S, lt, j, i, in); break;
    case PH_CORR_B: ph_corr_b<T>(C, S, lt, j, i, in, active); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Small utility kernels
// ---------------------------------------------------------------------------
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

}  // namespace pf

// ===========================================================================
// Host side
// ===========================================================================
using namespace pf;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(PF_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));           \
  } while (0)

struct pf_batch {
  int prec = 0;
  int max_cav = 0;
  int64_t max_tiles = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<PfCav> hcav;       // slot -> description (nx == 0: free)
  std::vector<char> used;
  PfCav* dcav = nullptr;
  PfScal* dscal = nullptr;
  int* dtile = nullptr;
  int* dactive = nullptr;
  PfScal* hscal = nullptr;  // pinned mirror
  int* hactive = nullptr;   // pinned
  int total_tiles = 0;
  bool tables_dirty = true;
  int64_t rounds = 0;
  int64_t predicted_rounds = 0;
  int64_t round_guard = 0;
  hipEvent_t ev = nullptr;
  bool chunk_pending = false;
};

static size_t esize(int prec) { return prec == PF_PREC_MIXED ? sizeof(float) : sizeof(double); }
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static int tiles_of(int nx, int ny, int* tx, int* ty) {
  *tx = (nx + TX - 1) / TX;
  *ty = (ny + TY - 1) / TY;
  return (*tx) * (*ty);
}

static int check_params(const pf_cavity_params* p) {
  if (!p) return fail(PF_EINVAL, "params is NULL");
  if (p->nx < 2 || p->ny < 2) return fail(PF_EINVAL, "nx and ny must be >= 2");
  if ((int64_t)p->nx * p->ny > (int64_t)1 << 31) return fail(PF_EINVAL, "cavity too large");
  if (p->scheme < 0 || p->scheme > 2) return fail(PF_EINVAL, "bad scheme");
  if (p->op < 0 || p->op > 1) return fail(PF_EINVAL, "bad op");
  if (p->mode < 0 || p->mode > 2) return fail(PF_EINVAL, "bad mode");
  return PF_OK;
}

static size_t cavity_bytes(int prec, const pf_cavity_params* p) {
  const size_t n = (size_t)p->nx * p->ny;
  int tx, ty;
  const size_t nt = tiles_of(p->nx, p->ny, &tx, &ty);
  size_t s = 0;
  s += 9 * align256(n * esize(prec));                          // u v p us vs r d0 d1 x
  s += align256(n * sizeof(double));                           // b
  if (prec == PF_PREC_MIXED) s += align256(n * sizeof(double));  // x64
  s += align256(nt * NP * sizeof(double));                     // partials
  return s;
}

static void fill_params(PfCav& c, const pf_cavity_params* p, int prec) {
  c.nx = p->nx;
  c.ny = p->ny;
  int tx, ty;
  c.ntiles = tiles_of(p->nx, p->ny, &tx, &ty);
  c.tiles_x = tx;
  c.scheme = p->scheme;
  c.lid_on = p->lid_enabled ? 1 : 0;
  c.jacobi = p->jacobi ? 1 : 0;
  c.maxiter = p->lin_maxiter;
  c.op = p->op;
  c.mode = p->mode;
  c.prec = prec;
  c.lid = p->lid_velocity;
  c.dx = p->dx;
  c.dy = p->dy;
  c.nu = p->nu;
  c.cfl = p->cfl_target;
  c.growth = p->cfl_growth;
  c.tol = p->lin_tol;
  c.invd = p->inv_diag;
  c.inner_rtol = (p->inner_rtol > 0.0) ? p->inner_rtol : 1e-4;
}

static int sync_tables(pf_batch* b) {
  if (!b->tables_dirty) return PF_OK;
  std::vector<int> tiles;
  tiles.reserve(b->max_tiles);
  for (int s = 0; s < b->max_cav; ++s) {
    if (!b->used[s]) continue;
    PfCav& c = b->hcav[s];
    c.tile0 = (int)tiles.size();
    for (int t = 0; t < c.ntiles; ++t) tiles.push_back(s);
  }
  if ((int64_t)tiles.size() > b->max_tiles) return fail(PF_EINVAL, "batch tile capacity exceeded");
  b->total_tiles = (int)tiles.size();
  HIPCHK(hipMemcpyAsync(b->dcav, b->hcav.data(), sizeof(PfCav) * b->max_cav, hipMemcpyHostToDevice, b->stream));
  if (!tiles.empty())
    HIPCHK(hipMemcpyAsync(b->dtile, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tables_dirty = false;
  return PF_OK;
}

static int push_scal(pf_batch* b, int slot) {
  HIPCHK(hipMemcpyAsync(b->dscal + slot, b->hscal + slot, sizeof(PfScal), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

static int pull_scal(pf_batch* b) {
  HIPCHK(hipMemcpyAsync(b->hscal, b->dscal, sizeof(PfScal) * b->max_cav, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

static int check_slot(pf_batch* b, int slot) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  if (slot < 0 || slot >= b->max_cav || !b->used[slot]) return fail(PF_EINVAL, "bad cavity slot");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  return PF_OK;
}

template <class T>
static void launch_round(pf_batch* b) {
  if (b->total_tiles <= 0) return;
  dim3 blk(TX, TY);
  k_round<T, 0><<<dim3(b->total_tiles), blk, 0, b->stream>>>(b->dcav, b->dscal, b->dtile, b->dactive);
}

static void launch_rounds(pf_batch* b, int64_t n) {
  for (int64_t r = 0; r < n; ++r) {
    if (b->prec == PF_PREC_MIXED)
      launch_round<float>(b);
    else
      launch_round<double>(b);
  }
  b->rounds += n;
}

static void fill_diag(const pf_batch* b, int slot, pf_step_diag* d) {
  const PfScal& s = b->hscal[slot];
  std::memset(d, 0, sizeof(*d));
  d->dt = s.dt;
  d->cfl = s.cfl_now;
  d->umax = s.umax;
  d->vmax = s.vmax;
  d->dt_prev = s.dt_prev;
  for (int k = 0; k < 5; ++k) d->hist[k] = s.hist[k];
  d->hist_len = s.hist_len;
  d->safeguard = s.safeguard;
  d->div_before = s.div_before;
  d->div_initial = s.div_initial;
  d->div_norm = s.div_after;
  d->rhs_mean = s.mean;
  d->bnorm = s.bnorm;
  d->rp_residual = s.rnorm_true;
  d->ru = s.ru;
  d->rv = s.rv;
  d->continuity = s.continuity;
  d->rp_info = s.info;
  d->rp_iterations = s.it_total;
  d->rp_converged = s.converged;
  d->nan_detected = s.nan;
  d->outer_passes = s.outer;
  d->steps_done = s.steps_done;
}

template <class T>
static int upload_field(pf_batch* b, PfCav& c, T* dst, const double* src, int64_t ld, int ghost) {
  const size_t n = (size_t)c.nx * c.ny;
  const double* s0 = src + (ghost ? (ld + 1) : 0);
  double* stage = c.b;  // b is recomputed every step; free between runs
  HIPCHK(hipMemcpy2DAsync(stage, sizeof(double) * c.nx, s0, sizeof(double) * ld, sizeof(double) * c.nx, c.ny,
                          hipMemcpyHostToDevice, b->stream));
  if (sizeof(T) == sizeof(double)) {
    HIPCHK(hipMemcpyAsync(dst, stage, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  } else {
    k_f64_to<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream>>>(stage, dst, n);
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

template <class T>
static int download_field(pf_batch* b, PfCav& c, const T* src, double* dst, int64_t ld, int ghost) {
  const size_t n = (size_t)c.nx * c.ny;
  double* d0 = dst + (ghost ? (ld + 1) : 0);
  double* stage = c.b;
  if (sizeof(T) == sizeof(double)) {
    stage = (double*)src;
  } else {
    k_to_f64<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream>>>(src, stage, n);
  }
  HIPCHK(hipMemcpy2DAsync(d0, sizeof(double) * ld, stage, sizeof(double) * c.nx, sizeof(double) * c.nx, c.ny,
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

const char* pf_last_error(void) { return g_err.c_str(); }

size_t pf_sizeof_params(void) { return sizeof(pf_cavity_params); }
size_t pf_sizeof_diag(void) { return sizeof(pf_step_diag); }

int pf_device_init(int device) {
  HIPCHK(hipSetDevice(device));
  return PF_OK;
}

int pf_batch_workspace_bytes(int precision, int ncav, const pf_cavity_params* params, size_t* bytes) {
  if (!bytes || ncav < 0 || (ncav > 0 && !params)) return fail(PF_EINVAL, "bad arguments");
  if (precision != PF_PREC_FP64 && precision != PF_PREC_MIXED) return fail(PF_EINVAL, "bad precision");
  size_t s = 0;
  for (int c = 0; c < ncav; ++c) {
    int rc = check_params(&params[c]);
    if (rc) return rc;
    s += cavity_bytes(precision, &params[c]);
  }
  *bytes = s;
  return PF_OK;
}

int pf_batch_create_empty(int precision, int max_cav, int64_t max_tiles, void* stream, pf_batch** out) {
  if (!out || max_cav <= 0 || max_tiles <= 0) return fail(PF_EINVAL, "bad arguments");
  if (precision != PF_PREC_FP64 && precision != PF_PREC_MIXED) return fail(PF_EINVAL, "bad precision");
  pf_batch* b = new pf_batch();
  b->prec = precision;
  b->max_cav = max_cav;
  b->max_tiles = max_tiles;
  b->stream = (hipStream_t)stream;
  b->hcav.assign(max_cav, PfCav{});
  b->used.assign(max_cav, 0);
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dcav, sizeof(PfCav) * max_cav);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dscal, sizeof(PfScal) * max_cav);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dtile, sizeof(int) * max_tiles);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dactive, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hscal, sizeof(PfScal) * max_cav, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hactive, sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    std::string msg = std::string("allocation failed: ") + hipGetErrorString(e);
    pf_batch_destroy(b);
    return fail(PF_ENOMEM, msg);
  }
  std::memset(b->hscal, 0, sizeof(PfScal) * max_cav);
  *b->hactive = 0;
  e = hipMemsetAsync(b->dscal, 0, sizeof(PfScal) * max_cav, b->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b->dtile, 0xff, sizeof(int) * max_tiles, b->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b->dactive, 0, sizeof(int), b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    std::string msg = std::string("init failed: ") + hipGetErrorString(e);
    pf_batch_destroy(b);
    return fail(PF_EHIP, msg);
  }
  *out = b;
  return PF_OK;
}

int pf_batch_add(pf_batch* b, const pf_cavity_params* params, void* workspace, size_t bytes, int* slot) {
  if (!b || !slot || !workspace) return fail(PF_EINVAL, "bad arguments");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  int rc = check_params(params);
  if (rc) return rc;
  if (((uintptr_t)workspace) & 255) return fail(PF_EINVAL, "workspace must be 256-byte aligned");
  const size_t need = cavity_bytes(b->prec, params);
  if (bytes < need) return fail(PF_EINVAL, "workspace too small");
  int s = -1;
  for (int k = 0; k < b->max_cav; ++k)
    if (!b->used[k]) {
      s = k;
      break;
    }
  if (s < 0) return fail(PF_ENOMEM, "no free cavity slot");
  PfCav c{};
  fill_params(c, params, b->prec);
  int64_t tiles_now = 0;
  for (int k = 0; k < b->max_cav; ++k)
    if (b->used[k]) tiles_now += b->hcav[k].ntiles;
  if (tiles_now + c.ntiles > b->max_tiles) return fail(PF_ENOMEM, "batch tile capacity exceeded");
  const size_t n = (size_t)c.nx * c.ny;
  const size_t vb = align256(n * esize(b->prec));
  char* w = (char*)workspace;
  void** vecs[9] = {&c.u, &c.v, &c.p, &c.us, &c.vs, &c.r, &c.d0, &c.d1, &c.x};
  for (int k = 0; k < 9; ++k) {
    *vecs[k] = w;
    w += vb;
  }
  c.b = (double*)w;
  w += align256(n * sizeof(double));
  if (b->prec == PF_PREC_MIXED) {
    c.x64 = (double*)w;
    w += align256(n * sizeof(double));
  } else {
    c.x64 = (double*)c.x;
  }
  c.part = (double*)w;
  HIPCHK(hipMemsetAsync(workspace, 0, need, b->stream));
  b->hcav[s] = c;
  b->used[s] = 1;
  PfScal& sc = b->hscal[s];
  std::memset(&sc, 0, sizeof(sc));
  sc.phase = PH_IDLE;
  sc.dt_prev = 1e-2;
  b->tables_dirty = true;
  rc = push_scal(b, s);
  if (rc) return rc;
  *slot = s;
  return PF_OK;
}

int pf_batch_remove(pf_batch* b, int slot) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  b->used[slot] = 0;
  b->hcav[slot] = PfCav{};
  std::memset(&b->hscal[slot], 0, sizeof(PfScal));
  b->tables_dirty = true;
  return push_scal(b, slot);
}

int pf_batch_create(int precision, int ncav, const pf_cavity_params* params, void* workspace,
                    size_t workspace_bytes, void* stream, pf_batch** out) {
  size_t need = 0;
  int rc = pf_batch_workspace_bytes(precision, ncav, params, &need);
  if (rc) return rc;
  if (ncav <= 0) return fail(PF_EINVAL, "ncav must be > 0");
  if (workspace_bytes < need) return fail(PF_EINVAL, "workspace too small");
  int64_t tiles = 0;
  for (int c = 0; c < ncav; ++c) {
    int tx, ty;
    tiles += tiles_of(params[c].nx, params[c].ny, &tx, &ty);
  }
  pf_batch* b = nullptr;
  rc = pf_batch_create_empty(precision, ncav, tiles, stream, &b);
  if (rc) return rc;
  char* w = (char*)workspace;
  for (int c = 0; c < ncav; ++c) {
    const size_t nb = cavity_bytes(precision, &params[c]);
    int slot = -1;
    rc = pf_batch_add(b, &params[c], w, nb, &slot);
    if (rc) {
      pf_batch_destroy(b);
      return rc;
    }
    w += nb;
  }
  rc = sync_tables(b);
  if (rc) {
    pf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PF_OK;
}

int pf_batch_destroy(pf_batch* b) {
  if (!b) return PF_OK;
  (void)hipStreamSynchronize(b->stream);
  if (b->dcav) (void)hipFree(b->dcav);
  if (b->dscal) (void)hipFree(b->dscal);
  if (b->dtile) (void)hipFree(b->dtile);
  if (b->dactive) (void)hipFree(b->dactive);
  if (b->hscal) (void)hipHostFree(b->hscal);
  if (b->hactive) (void)hipHostFree(b->hactive);
  if (b->ev) (void)hipEventDestroy(b->ev);
  delete b;
  return PF_OK;
}

int pf_batch_set_params(pf_batch* b, int slot, const pf_cavity_params* params) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = check_params(params);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (params->nx != c.nx || params->ny != c.ny) return fail(PF_EINVAL, "nx/ny cannot change");
  fill_params(c, params, b->prec);
  b->tables_dirty = true;
  return PF_OK;
}

int pf_batch_upload(pf_batch* b, int slot, const double* u, const double* v, const double* p, int64_t ld,
                    int ghost) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (ld < c.nx + (ghost ? 2 : 0)) return fail(PF_EINVAL, "ld too small");
  const void* dsts[3] = {c.u, c.v, c.p};
  const double* srcs[3] = {u, v, p};
  for (int k = 0; k < 3; ++k) {
    if (!srcs[k]) continue;
    if (b->prec == PF_PREC_MIXED)
      rc = upload_field<float>(b, c, (float*)dsts[k], srcs[k], ld, ghost);
    else
      rc = upload_field<double>(b, c, (double*)dsts[k], srcs[k], ld, ghost);
    if (rc) return rc;
  }
  return PF_OK;
}

int pf_batch_download(pf_batch* b, int slot, double* u, double* v, double* p, int64_t ld, int ghost) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (ld < c.nx + (ghost ? 2 : 0)) return fail(PF_EINVAL, "ld too small");
  const void* srcs[3] = {c.u, c.v, c.p};
  double* dsts[3] = {u, v, p};
  for (int k = 0; k < 3; ++k) {
    if (!dsts[k]) continue;
    if (b->prec == PF_PREC_MIXED)
      rc = download_field<float>(b, c, (const float*)srcs[k], dsts[k], ld, ghost);
    else
      rc = download_field<double>(b, c, (const double*)srcs[k], dsts[k], ld, ghost);
    if (rc) return rc;
  }
  return PF_OK;
}

int pf_batch_zero_fields(pf_batch* b) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  for (int s = 0; s < b->max_cav; ++s) {
    if (!b->used[s]) continue;
    PfCav& c = b->hcav[s];
    const size_t n = (size_t)c.nx * c.ny * esize(b->prec);
    HIPCHK(hipMemsetAsync(c.u, 0, n, b->stream));
    HIPCHK(hipMemsetAsync(c.v, 0, n, b->stream));
    HIPCHK(hipMemsetAsync(c.p, 0, n, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int pf_batch_set_meta(pf_batch* b, int slot, double dt_prev, const double* hist, int hist_len) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (hist_len < 0 || (hist_len > 0 && !hist)) return fail(PF_EINVAL, "bad history");
  PfScal& s = b->hscal[slot];
  s.dt_prev = dt_prev;
  // keep the last 5 entries, as hist[-5:] does
  const int n = hist_len > 5 ? 5 : hist_len;
  for (int k = 0; k < 5; ++k) s.hist[k] = 0.0;
  for (int k = 0; k < n; ++k) s.hist[k] = hist[hist_len - n + k];
  s.hist_len = n;
  return push_scal(b, slot);
}

int pf_batch_set_dt(pf_batch* b, int slot, double dt) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  b->hscal[slot].dt = dt;
  return push_scal(b, slot);
}

int pf_batch_set_rhs(pf_batch* b, int slot, const double* rhs, int64_t ld) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (!rhs || ld < c.nx) return fail(PF_EINVAL, "bad rhs");
  HIPCHK(hipMemcpy2DAsync(c.b, sizeof(double) * c.nx, rhs, sizeof(double) * ld, sizeof(double) * c.nx, c.ny,
                          hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int pf_batch_get_solution(pf_batch* b, int slot, double* x, int64_t ld) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (!x || ld < c.nx) return fail(PF_EINVAL, "bad output");
  const double* src = b->hscal[slot].x_is_b ? c.b : c.x64;
  HIPCHK(hipMemcpy2DAsync(x, sizeof(double) * ld, src, sizeof(double) * c.nx, sizeof(double) * c.nx, c.ny,
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

static int arm(pf_batch* b, int nsteps, const int32_t* active) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  if (nsteps <= 0) return fail(PF_EINVAL, "nsteps must be > 0");
  int rc = sync_tables(b);
  if (rc) return rc;
  int nact = 0;
  int64_t guard = 64;
  for (int s = 0; s < b->max_cav; ++s) {
    PfScal& sc = b->hscal[s];
    if (!b->used[s]) continue;
    const bool on = !active || active[s];
    sc.ticket = 0;
    if (!on) {
      sc.phase = PH_IDLE;
      continue;
    }
    const PfCav& c = b->hcav[s];
    sc.steps_done = 0;
    if (c.mode == PF_MODE_STEP) {
      sc.phase = PH_STATS;
      sc.steps_left = nsteps;
    } else if (c.mode == PF_MODE_PROJECT) {
      sc.phase = PH_PRED_A;
      sc.steps_left = 1;
    } else {
      sc.phase = PH_PINIT;
      sc.steps_left = 1;
    }
    const int64_t per = 4 * (int64_t)std::max(c.maxiter, 0) + 64;
    guard = std::max(guard, (int64_t)sc.steps_left * per + 64);
    ++nact;
  }
  *b->hactive = nact;
  HIPCHK(hipMemcpyAsync(b->dscal, b->hscal, sizeof(PfScal) * b->max_cav, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->dactive, b->hactive, sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->rounds = 0;
  b->round_guard = guard;
  return PF_OK;
}

static int chunk(pf_batch* b, int64_t n) {
  launch_rounds(b, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->hactive, b->dactive, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipEventRecord(b->ev, b->stream));
  b->chunk_pending = true;
  return PF_OK;
}

int pf_batch_run(pf_batch* b, int nsteps, const int32_t* active, pf_step_diag* diag) {
  int rc = arm(b, nsteps, active);
  if (rc) return rc;
  if (*b->hactive > 0) {
    int64_t first = std::max<int64_t>(8, (b->predicted_rounds * 15) / 16);
    int64_t step = 16;
    rc = chunk(b, first);
    if (rc) return rc;
    while (true) {
      HIPCHK(hipEventSynchronize(b->ev));
      b->chunk_pending = false;
      if (*b->hactive <= 0) break;
      if (b->rounds > b->round_guard) return fail(PF_ESTATE, "round limit exceeded (internal error)");
      rc = chunk(b, step);
      if (rc) return rc;
      step = std::min<int64_t>(step * 2, 256);
    }
    b->predicted_rounds = b->rounds;
  }
  rc = pull_scal(b);
  if (rc) return rc;
  if (diag)
    for (int s = 0; s < b->max_cav; ++s)
      if (b->used[s]) fill_diag(b, s, &diag[s]);
  return PF_OK;
}

int pf_batch_launch(pf_batch* b, int nsteps, const int32_t* active) {
  int rc = arm(b, nsteps, active);
  if (rc) return rc;
  if (*b->hactive <= 0) return PF_OK;
  return chunk(b, 16);
}

int pf_batch_poll(pf_batch* b, int* done) {
  if (!b || !done) return fail(PF_EINVAL, "bad arguments");
  if (!b->chunk_pending) {
    *done = 1;
    return PF_OK;
  }
  hipError_t q = hipEventQuery(b->ev);
  if (q == hipErrorNotReady) {
    *done = 0;
    return PF_OK;
  }
  HIPCHK(q);
  b->chunk_pending = false;
  if (*b->hactive <= 0) {
    int rc = pull_scal(b);
    if (rc) return rc;
    *done = 1;
    return PF_OK;
  }
  if (b->rounds > b->round_guard) return fail(PF_ESTATE, "round limit exceeded (internal error)");
  int rc = chunk(b, 32);
  if (rc) return rc;
  *done = 0;
  return PF_OK;
}

int pf_batch_diag(pf_batch* b, pf_step_diag* diag) {
  if (!b || !diag) return fail(PF_EINVAL, "bad arguments");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  for (int s = 0; s < b->max_cav; ++s)
    if (b->used[s]) fill_diag(b, s, &diag[s]);
  return PF_OK;
}

int pf_batch_centerline(pf_batch* b, int slot, double* out) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (!out) return fail(PF_EINVAL, "out is NULL");
  const int cx = c.nx / 2;
  const size_t es = esize(b->prec);
  std::vector<char> tmp((size_t)c.ny * es);
  HIPCHK(hipMemcpy2DAsync(tmp.data(), es, (const char*)c.u + cx * es, es * c.nx, es, c.ny, hipMemcpyDeviceToHost,
                          b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int j = 0; j < c.ny; ++j)
    out[j] = (es == sizeof(float)) ? (double)((const float*)tmp.data())[j] : ((const double*)tmp.data())[j];
  return PF_OK;
}

int pf_batch_rounds(pf_batch* b, int64_t* rounds) {
  if (!b || !rounds) return fail(PF_EINVAL, "bad arguments");
  *rounds = b->rounds;
  return PF_OK;
}

int pf_op_apply(int opk, int nx, int ny, double dx, double dy, const double* in0, const double* in1,
                const double* in2, double* out0, double* out1, void* stream) {
  if (nx < 1 || ny < 1) return fail(PF_EINVAL, "bad shape");
  const bool need_uv = (opk == PF_OPK_DIV_COHERENT || opk == PF_OPK_DIV_CENTRAL || opk == PF_OPK_ADV_QUICK ||
                        opk == PF_OPK_ADV_UPWIND);
  if (opk < 0 || opk > 7) return fail(PF_EINVAL, "bad operator");
  if (!in0 || !out0 || (need_uv && !in1) || ((opk == 4 || opk == 5) && !in2) ||
      (opk == PF_OPK_GRAD_COHERENT && !out1))
    return fail(PF_EINVAL, "missing buffer");
  if (opk == PF_OPK_DIV_CENTRAL && (nx < 2 || ny < 2)) return fail(PF_EINVAL, "central divergence needs nx,ny >= 2");
  dim3 grid((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_op<<<grid, dim3(TX, TY), 0, (hipStream_t)stream>>>(opk, nx, ny, dx, dy, in0, in1, in2, out0, out1);
  HIPCHK(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
