// Device phase machine of the projection step: per-cell phase bodies, the
// scalar "tails" that consume a phase's reduced partials, and the round kernel.
//
// Cross-workgroup protocol (no fences, no spin-waits)
// -------------------------------------------------
// Cavity c is cut into nreg regions of tpb 64x16-cell tiles stacked in y and
// covered by nblk workgroups; workgroup b owns regions [b*rpw, (b+1)*rpw)
// (one cell per thread per tile). In round r every workgroup of c
//   1. copies the cavity's scalar state written in round r-1 (buffer r^1),
//   2. if that state has a pending phase, sums the pending phase's FP64
//      partials of all nblk workgroups (written in round r-1, parity r^1) in
//      a fixed order and runs the phase's scalar tail (dt controller, CG
//      alpha/beta, convergence tests, phase transition) -- every workgroup
//      computes the same bits, so they all agree on the next phase,
//   3. executes that phase on its regions and writes its own partials (parity r),
//   4. (workgroup 0 only) stores the state, with the executed phase pending,
//      into buffer r.
// Data written in round r is only read in round r+1, so kernel boundaries are
// the only synchronisation, and all reductions are bitwise deterministic.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pf_cells.hpp"
#include "pyflow_hip.h"

namespace pf {

// Stores of the PCG rounds' r, p and x: streaming (non-temporal) by default;
// build with -DPF_STREAM_STORES=0 for ordinary stores.
#ifndef PF_STREAM_STORES
#define PF_STREAM_STORES 1
#endif
template <class V>
__device__ __attribute__((always_inline)) inline void pcg_store(V v, V* p) {
#if PF_STREAM_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

constexpr int TX = 64;        // tile width = wavefront width
constexpr int TY = 16;        // tile height; one wave per tile row
constexpr int NTH = TX * TY;  // threads per workgroup
constexpr int NWAVE = NTH / 64;
constexpr int NP = 16;        // FP64 partial values a thread carries (register arrays)
constexpr int NPS = 32;       // FP64 partial slots per workgroup in memory (stride)
constexpr int KMAX = 3;       // slots of every phase except the PCG rounds below
constexpr int KXD = 6;        // slots of the FP64-iterate PCG round
constexpr int KCA = 14;       // slots of the two-iteration PCG round
constexpr int KCA3 = 27;      // slots of the three-iteration PCG round

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
  PH_CG = 13,   // MIXED: one single-reduction (Chronopoulos-Gear) PCG iteration, or two
                // s-step iterations on a cavity with PfCav::ca
  PH_DEFL_A = 14,  // MIXED deflation: y.(b - A x) of the finished solve
  PH_DEFL_B = 15,  // MIXED deflation: x += coef y
  PH_DSTORE = 16,  // MIXED deflation set-up: y = float(x) of the solve of A y = 1, and y.A y
  PH_DONE = 17,
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
    case PH_DEFL_A: return 1;
    case PH_DSTORE: return 1;
    default: return 0;
  }
}
__device__ __host__ inline unsigned phase_maxmask(int ph) {
  return ph == PH_STATS ? 0x3u : (ph == PH_CORR_A ? 0x4u : 0u);
}

// Read-mostly description of one cavity (rewritten only between runs).
struct PfCav {
  int nx, ny, tiles_x, tiles_y, ntiles, nblk, blk0, tpb;  // nblk workgroups; tpb = M: tiles per region, stacked in y
  int nreg, rpw;  // regions; regions per workgroup (workgroup b owns regions b*rpw .. b*rpw+rpw-1)
  int scheme, lid_on, jacobi, maxiter, op, mode, prec;
  int xd;  // MIXED: the PCG iterate is accumulated in FP64 (x64) instead of FP32
  int ca;  // MIXED s-step rounds: 0 one PCG iteration per round; 1 two (4-cell halo); 2 three (6-cell halo)
  float cxf, cyf, invdf;  // (float)(1/dx^2), (float)(1/dy^2), (float)invd, formed on the host
  double lid, dx, dy, nu, cfl, growth, tol, invd, inner_rtol;
  void *u, *v, *p, *us, *vs, *r, *d0, *d1, *x;  // element type T
  void* r1;                                     // MIXED: r double-buffered (r, r1); p in d0/d1
  double *b, *x64;                              // x64 aliases x in FP64 mode
  double* part;                                 // [2][NPS][nblk]: slot k of workgroup b at k*nblk + b
  float* y;                                     // MIXED: deflation vector (nullptr in FP64)
  int defl;  // MIXED, pinned operator: Galerkin correction along y = A^-1 1 after each solve
  int local;  // FP64, one region: the CG iterations of a solve run inside one round (local_cg)
};

// Scalar state of one cavity; two copies (round parity) per cavity.
struct PfScal {
  int phase;     // phase to execute next
  int pending;   // phase whose partials await their tail (0 = none)
  int steps_left, steps_done, it, it_total, info, outer, x_is_b;
  int hist_len, safeguard, nan, converged, it_run, outer_run, done_now;
  int ca_inc;      // two-iteration round: iterations the pending round runs (0: basis only)
  int rounds_run;  // rounds this cavity executed a phase in since it was armed (deflation set-up excluded)
  int dsetup;      // the running solve is the deflation set-up, A y = 1
  int dready;      // y holds the deflation vector of this cavity's operator
  int step_round;  // rounds_run when the last projection step ended
  int step_len;    // rounds_run of the last projection step (the host sizes its chunks with it)
  int pad_[2];
  double dden;     // y.A y of the stored y
  double dcoef;    // Galerkin coefficient of the finished solve
  double dt_prev, dt, cfl_now, umax, vmax, hist[5], div_before;
  double mean, div_initial, bnorm, atol, inner_atol, rho, rho_prev, beta, pq, alpha, rnorm2, rn_outer;
  double rnorm_true, div_after, ru, rv, continuity;
  double ca0, cb0, ca1, cb1;  // two-iteration round: alpha_k, beta_k, alpha_k+1, beta_k+1
  double ca2, cb2;            // three-iteration round: alpha_k+2, beta_k+2
};
static_assert(sizeof(PfScal) % 4 == 0, "PfScal copied as 32-bit words");
// The round kernels keep a PfScal in LDS ahead of their reduction arrays; a
// size that is a multiple of 16 bytes keeps those arrays 16-byte aligned, so
// their accesses compile to 128-bit LDS operations.
static_assert(sizeof(PfScal) % 16 == 0, "reduction arrays after the LDS state stay 16-byte aligned");

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

// Wave reduction of the first K of NP values (bit k of MM set: NaN-propagating
// maximum, else sum); the totals end in lane 63.
template <int K, unsigned MM>
__device__ inline void wave_reduce_k(double (&v)[NP]) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((MM >> k) & 1u) ? wave_reduce<true>(v[k]) : wave_reduce<false>(v[k]);
}

// Partials are stored slot-major: slot k of workgroup b at part[k * nblk + b]
// (per round parity), so a wave that folds one slot over all workgroups reads
// consecutive doubles. A writer gets dst = &part[b] and stride = nblk.
//
// Workgroup partials of the first K of NP per-thread values: every wave
// reduces on the DPP crossbar, lane 63 parks the wave total in sm[wave], and
// after one barrier thread k < K folds the NWAVE totals in wave order and
// stores slot k to dst[k * stride].
template <int K, unsigned MM>
__device__ inline void block_partials_k(double (&v)[NP], double (*sm)[NPS], double* dst, int stride) {
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
  wave_reduce_k<K, MM>(v);
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[wave][k] = v[k];
  }
  __syncthreads();
  if (tid < K) {
    const bool mx = (MM >> tid) & 1u;
    double t = sm[0][tid];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) t = mx ? nanmax(t, sm[w][tid]) : t + sm[w][tid];
    dst[(size_t)tid * stride] = t;
  }
}

// Dispatch on a phase's partial layout (compile-time K and reduction ops).
#define PF_PHASE_LAYOUT(F)                  \
  switch (ph) {                             \
    case PH_STATS: F(3, 0x3u); break;       \
    case PH_CORR_A: F(3, 0x4u); break;      \
    case PH_PINIT:                          \
    case PH_OUTER_B:                        \
    case PH_CG0:                            \
    case PH_CG: F(3, 0u); break;            \
    case PH_PRED_B:                         \
    case PH_Q:                              \
    case PH_CORR_B: F(2, 0u); break;        \
    case PH_P:                              \
    case PH_DEFL_A:                         \
    case PH_DSTORE:                         \
    case PH_FIN: F(1, 0u); break;           \
    default: break;                         \
  }

__device__ inline void block_partials_phase(int ph, double (&v)[NP], double (*sm)[NPS], double* dst, int stride) {
#define PF_BP(K, MM) block_partials_k<K, MM>(v, sm, dst, stride)
  PF_PHASE_LAYOUT(PF_BP)
#undef PF_BP
}
__device__ inline void wave_reduce_phase(int ph, double (&v)[NP]) {
#define PF_WR(K, MM) wave_reduce_k<K, MM>(v)
  PF_PHASE_LAYOUT(PF_WR)
#undef PF_WR
}

// ---------------------------------------------------------------------------
// Scalar tails: run by one thread on the workgroup's shared copy of the state
// after the pending phase's partials have been summed into tot[].
// ---------------------------------------------------------------------------
// Deflation (MIXED, pinned operator). The pinned Neumann operator has one
// small isolated eigenvalue; its eigenvector is smooth, nearly constant away
// from the pinned cell. FP32 rounding in the inner passes leaves an error
// along it whose residual lies far below the tolerance, so no later pass
// removes it, and the pressure carries a near-constant offset (2.7e-4
// relative after the first step of the 1024^2 cavity, against 2.2e-6 for the
// FP64 recurrence). After every solve the cavity takes the Galerkin step
// x += y (y.r)/(y.A y) along y = A^-1 1, which minimises the A-norm error
// along y and removes that offset. y is solved once per cavity, to kDeflTol:
// the step is exact for whatever y is stored, and a y solved to 1e-1 gave the
// same correction as one solved to 1e-12 on that cavity.
constexpr double kDeflTol = 1e-2;

// Enter PINIT. A deflating cavity without its y first solves A y = 1 (the
// set-up, flagged in dsetup from its PINIT round on); PH_DSTORE stores y and
// enters PINIT again for the solve the step needs.
__device__ __host__ inline void enter_pinit(const PfCav& C, PfScal& S) {
  S.dsetup = (C.defl && !S.dready) ? 1 : 0;
  S.phase = PH_PINIT;
}

// First phase of a refinement pass: with an FP64 iterate there is no FP32
// correction to fold in, so the pass starts with the FP64 residual.
__device__ __attribute__((always_inline)) inline int outer_entry(const PfCav& C) { return C.xd ? PH_OUTER_B : PH_OUTER_A; }

__device__ __attribute__((always_inline)) inline void cg_next(const PfCav& C, PfScal& S) {
  const bool mixed = (C.prec == PF_PREC_MIXED);
  if (S.it_total >= C.maxiter) {
    S.info = C.maxiter;
    S.phase = mixed ? outer_entry(C) : PH_FIN;
    return;
  }
  if (sqrt(S.rnorm2) < S.inner_atol) {
    if (!mixed) S.info = 0;
    S.phase = mixed ? outer_entry(C) : PH_FIN;
    return;
  }
  S.phase = mixed ? PH_CG : PH_P;
}

// First round of an inner FP32 solve: CG0, or on a two-iteration cavity a
// round that runs no iteration and only returns the partials of the starting
// basis (p = M r).
__device__ inline void inner_start(const PfCav& C, PfScal& S) {
  if (C.ca) {
    S.phase = PH_CG;
    S.ca_inc = 0;
    S.ca0 = S.cb0 = S.ca1 = S.cb1 = S.ca2 = S.cb2 = 0.0;
  } else {
    S.phase = PH_CG0;
  }
}

__device__ inline void enter_done(PfScal& S) {
  S.phase = PH_DONE;
  S.done_now = 1;
}

// Tail of a MIXED PCG iteration: beta = gamma'/gamma, alpha' = gamma'/(delta' -
// beta gamma'/alpha). Pending phase PH_CG is the common case of a round.
__device__ __attribute__((always_inline)) inline void tail_cg(const PfCav& C, PfScal& S, const double* tot) {
  S.pending = PH_IDLE;
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
}

__attribute__((noinline)) // Tail of an FP64-iterate PCG iteration. Besides gamma' = r'.u', delta' =
// (A u').u' and |r'|^2 the round returned pi = p.(A p), sigma = u'.(A p) and
// rho = r'.p of the direction p just used, so p' = u' + beta p has
//   p'.(A p') = delta' + 2 beta sigma + beta^2 pi   (A is symmetric on the
//   vectors the solve produces: their pinned cell stays 0)
// and the step along p' is the exact line minimum alpha' = r'.p' / p'.(A p')
// with r'.p' = gamma' + beta rho. No orthogonality between successive
// residuals is assumed, which the FP32 vectors lose on large grids.
__device__ __attribute__((always_inline)) inline void tail_cg_xd(const PfCav& C, PfScal& S, const double* tot) {
  S.pending = PH_IDLE;
  const double g1 = tot[0], d1 = tot[1], pi = tot[3], sig = tot[4], rho = tot[5];
  const double beta = g1 / S.rho;
  const double pap = d1 + 2.0 * beta * sig + beta * beta * pi;
  S.alpha = (g1 + beta * rho) / pap;
  S.beta = beta;
  S.rho = g1;
  S.rnorm2 = tot[2];
  S.it += 1;
  S.it_total += 1;
  S.it_run += 1;
  cg_next(C, S);
}

// Tail of a two-iteration PCG round (s-step PCG with s = 2: Chronopoulos &
// Gear 1989; in the communication-avoiding form of Hoemmen 2010, ch. 5, and
// Carson 2015). The round returned, for the new r, p and z = M r, the Gram
// matrix W = Y^T M^-1 Y of the basis Y = [p, Bp, B^2 p, z, Bz], B = M A,
// (13 entries: W01 W02 W04 W11 W12 W13 W14 W22 W23 W24 W33 W34 W44) and |r|^2.
// Every quantity of the next two PCG iterations is a quadratic form in W:
// with e_i the coordinate vectors of Y and T the shift (B Y e_i = Y T e_i):
//   gamma_k = W33,                  alpha_k   = gamma_k / (e0.W.T e0)
//   z_k+1 = e3 - alpha_k e1,        gamma_k+1 = z_k+1.W.z_k+1,  beta_k = gamma_k+1/gamma_k
//   p_k+1 = z_k+1 + beta_k e0,      alpha_k+1 = gamma_k+1 / (p_k+1.W.T p_k+1)
//   z_k+2 = z_k+1 - alpha_k+1 T p_k+1,  beta_k+1 = z_k+2.W.z_k+2 / gamma_k+1
// W is a weighted Gram matrix, symmetric whatever A is, so the identity rows
// of the pinned and Dirichlet operators need no special case.
__device__ __attribute__((always_inline)) inline void tail_ca(const PfCav& C, PfScal& S, const double* t) {
  S.pending = PH_IDLE;
  const int inc = S.ca_inc;
  S.it += inc;
  S.it_total += inc;
  S.it_run += inc;
  S.rnorm2 = t[13];
  S.rho = t[10];
  if (S.it_total >= C.maxiter) {
    S.info = C.maxiter;
    S.phase = outer_entry(C);
    return;
  }
  if (sqrt(S.rnorm2) < S.inner_atol) {
    S.phase = outer_entry(C);
    return;
  }
  const double w01 = t[0], w02 = t[1], w04 = t[2], w11 = t[3], w12 = t[4], w13 = t[5], w14 = t[6];
  const double w22 = t[7], w23 = t[8], w24 = t[9], w33 = t[10], w34 = t[11], w44 = t[12];
  double a0 = w33 / w01;
  double b0 = 0.0, a1 = 0.0, b1 = 0.0;
  int n = 1;
  // alpha < 0 is legitimate: the Dirichlet operator (+lap) is negative
  // definite. gamma = r.M r > 0 whatever the sign of A.
  if (a0 == 0.0 || !isfinite(a0)) {
    // No usable step (only reachable with corrupted input): count one
    // iteration and restart the direction at p = M r.
    a0 = 0.0;
  } else {
    const double g1 = w33 - 2.0 * a0 * w13 + a0 * a0 * w11;
    b0 = g1 / w33;
    const double pap = (w34 - a0 * w23 + b0 * w13) + (-a0 * w14 + a0 * a0 * w12 - a0 * b0 * w11) +
                       (b0 * w04 - a0 * b0 * w02 + b0 * b0 * w01);
    const double a1c = g1 / pap;
    if (C.maxiter - S.it_total >= 2 && g1 > 0.0 && pap != 0.0 && isfinite(a1c)) {
      a1 = a1c;
      n = 2;
      const double k1 = -(a0 + a1 * b0), k2 = a1 * a0, k4 = -a1;
      const double g2 = w33 + k1 * k1 * w11 + k2 * k2 * w22 + k4 * k4 * w44 +
                        2.0 * (k1 * w13 + k2 * w23 + k4 * w34 + k1 * k2 * w12 + k1 * k4 * w14 + k2 * k4 * w24);
      b1 = g2 / g1;
      if (!isfinite(b1)) b1 = 0.0;
    }
    if (!isfinite(b0)) b0 = 0.0;
  }
  S.ca0 = a0;
  S.cb0 = b0;
  S.ca1 = a1;
  S.cb1 = b1;
  S.ca_inc = n;
  S.alpha = n == 2 ? a1 : a0;
  S.beta = n == 2 ? b1 : b0;
  S.phase = PH_CG;
}

// Tail of a three-iteration round (s = 3). The pending round's basis is
// Y = [p, Bp, B^2 p, B^3 p, z, Bz, B^2 z] (B = M A); t holds the entries of
// W = Y^T M^-1 Y that the recurrence reads (all but W00 and W04, slot
// w3_slot(i, j)) and |r|^2 (slot 26). In the coefficient space of Y, with T
// the shift along each chain (B Y c = Y T c), three PCG iterations are
//   delta_j = p_j.W.T p_j,  alpha_j = gamma_j / delta_j,
//   z_j+1 = z_j - alpha_j T p_j,  gamma_j+1 = z_j+1.W.z_j+1,  beta_j = gamma_j+1 / gamma_j,
//   p_j+1 = z_j+1 + beta_j p_j,
// from p_0 = e0, z_0 = e4, gamma_0 = W44. The supports of p_j, T p_j and
// z_j+1 are fixed (bit masks below), so each form sums only its nonzero
// terms. An iteration that the budget leaves out, or whose step breaks down,
// keeps alpha = beta = 0: the round then moves by the iterations before it
// and restarts the direction, as the two-iteration round does.
__host__ __device__ constexpr int w3_slot(int i, int j) {
  return (i > j) ? w3_slot(j, i)
                 : (i == 0 ? ((j == 0 || j == 4) ? -1 : (j < 4 ? j - 1 : j - 2))
                           : (i == 1 ? 5 : i == 2 ? 11 : i == 3 ? 16 : i == 4 ? 20 : i == 5 ? 23 : 25) + (j - i));
}
static_assert(w3_slot(6, 6) == 25 && w3_slot(0, 6) == 4 && w3_slot(3, 4) == 17 && w3_slot(5, 1) == 9, "W slots");
__host__ __device__ constexpr unsigned w3_shift_mask(unsigned m) { return ((m & 0x7u) << 1) | ((m & 0x30u) << 1); }

// u.W.v over the supports SU, SV, by template recursion so that every slot
// index is a compile-time constant (the tail reads tot[] at fixed offsets).
template <unsigned SV, int I, int J = 0>
__device__ __attribute__((always_inline)) inline double w3_row(const double* t, const double (&v)[7]) {
  if constexpr (J == 7) {
    return 0.0;
  } else if constexpr (((SV >> J) & 1u) != 0) {
    constexpr int k = w3_slot(I, J);
    static_assert(k >= 0, "the form reads only formed entries of W");
    return t[k] * v[J] + w3_row<SV, I, J + 1>(t, v);
  } else {
    return w3_row<SV, I, J + 1>(t, v);
  }
}
template <unsigned SU, unsigned SV, int I = 0>
__device__ __attribute__((always_inline)) inline double w3_form(const double* t, const double (&u)[7],
                                                                const double (&v)[7]) {
  if constexpr (I == 7) {
    return 0.0;
  } else if constexpr (((SU >> I) & 1u) != 0) {
    return u[I] * w3_row<SV, I>(t, v) + w3_form<SU, SV, I + 1>(t, u, v);
  } else {
    return w3_form<SU, SV, I + 1>(t, u, v);
  }
}

// One iteration in coefficient space; SP: support of p_j, SZ: of z_j+1.
// Returns whether a further iteration may follow.
template <unsigned SP, unsigned SZ>
__device__ __attribute__((always_inline)) inline bool ca3_iter(const double* t, double (&p)[7], double (&z)[7],
                                                               double& gam, double& al, double& be) {
  static_assert((SP & 0x48u) == 0, "the basis has no B^4 p or B^3 z");
  constexpr unsigned ST = w3_shift_mask(SP);
  static_assert((SZ & ST) == ST, "T p_j lies in the support of z_j+1");
  double tp[7] = {0.0, p[0], p[1], p[2], 0.0, p[4], p[5]};
  const double a = gam / w3_form<SP, ST>(t, p, tp);
  if (!(a != 0.0 && isfinite(a))) return false;
  double zn[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) zn[i] = z[i] - a * tp[i];
  const double g1 = w3_form<SZ, SZ>(t, zn, zn);
  double b = g1 / gam;
  if (!isfinite(b)) b = 0.0;
  al = a;
  be = b;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    p[i] = zn[i] + b * p[i];
    z[i] = zn[i];
  }
  gam = g1;
  return g1 > 0.0;
}

__device__ __attribute__((always_inline)) inline void tail_ca3(const PfCav& C, PfScal& S, const double* t) {
  S.pending = PH_IDLE;
  const int inc = S.ca_inc;
  S.it += inc;
  S.it_total += inc;
  S.it_run += inc;
  S.rnorm2 = t[26];
  S.rho = t[w3_slot(4, 4)];
  if (S.it_total >= C.maxiter) {
    S.info = C.maxiter;
    S.phase = outer_entry(C);
    return;
  }
  if (sqrt(S.rnorm2) < S.inner_atol) {
    S.phase = outer_entry(C);
    return;
  }
  double al[3] = {0.0, 0.0, 0.0}, be[3] = {0.0, 0.0, 0.0};
  double p[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double z[7] = {0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
  double gam = t[w3_slot(4, 4)];
  const int budget = C.maxiter - S.it_total;
  // A failed first step counts one iteration with alpha = 0 (no move).
  int n = 1;
  if (ca3_iter<0x01u, 0x12u>(t, p, z, gam, al[0], be[0]) && budget >= 2) {
    const bool more = ca3_iter<0x13u, 0x36u>(t, p, z, gam, al[1], be[1]);
    if (al[1] != 0.0) n = 2;
    if (more && n == 2 && budget >= 3) {
      ca3_iter<0x37u, 0x7Eu>(t, p, z, gam, al[2], be[2]);
      if (al[2] != 0.0) n = 3;
    }
  }
  S.ca0 = al[0];
  S.cb0 = be[0];
  S.ca1 = al[1];
  S.cb1 = be[1];
  S.ca2 = al[2];
  S.cb2 = be[2];
  S.ca_inc = n;
  S.alpha = al[n - 1];
  S.beta = be[n - 1];
  S.phase = PH_CG;
}

// Tails of the FP64 CG phases P and Q, shared by the generic tail and the
// in-round CG (local_cg); Q's phase transition is cg_next's.
__device__ __attribute__((always_inline)) inline void tail_p(PfScal& S, const double* tot) {
  S.pq = tot[0];
  S.alpha = S.rho / tot[0];
  S.phase = PH_Q;
}
__device__ __attribute__((always_inline)) inline void tail_q_sums(PfScal& S, const double* tot) {
  S.rho_prev = S.rho;
  S.rho = tot[0];
  S.rnorm2 = tot[1];
  S.beta = S.rho / S.rho_prev;
  S.it += 1;
  S.it_total += 1;
  S.it_run += 1;
}

// V tells instances apart: each kernel instance that calls the cold functions
// gets its own copies, compiled within that kernel's register budget.
template <int V>
__attribute__((noinline)) __device__ void tail(const PfCav& C, PfScal& S, const double* tot) {
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
      enter_pinit(C, S);
      break;
    case PH_PINIT: {
      const double bn = sqrt(tot[0]);
      S.bnorm = bn;
      S.atol = pymax(0.0, (S.dsetup ? kDeflTol : C.tol) * bn);
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
        S.phase = S.dsetup ? PH_DSTORE : PH_FIN;
        break;
      }
      if (C.prec == PF_PREC_MIXED) {
        if (bn < S.atol) {
          S.phase = PH_FIN;
          break;
        }
        S.inner_atol = pymax(C.inner_rtol * bn, 0.5 * S.atol);
        inner_start(C, S);
        break;
      }
      S.inner_atol = S.atol;
      S.phase = (sqrt(S.rnorm2) < S.atol) ? PH_FIN : PH_P;
      break;
    }
    case PH_P: tail_p(S, tot); break;
    case PH_Q:
      tail_q_sums(S, tot);
      cg_next(C, S);
      break;
    case PH_OUTER_A: S.phase = PH_OUTER_B; break;
    case PH_OUTER_B: {
      const double rn = sqrt(tot[0]);
      S.rn_outer = rn;
      S.outer += 1;
      S.outer_run += 1;
      // The solve ends when the FP64 residual meets the tolerance, also when
      // the iteration budget ran out on this iterate (it then met the
      // stopping rule within lin_maxiter iterations), or when the budget is
      // spent.
      if (rn < S.atol) S.info = 0;
      if (rn < S.atol || S.info != 0) {
        if (S.dsetup)
          S.phase = PH_DSTORE;
        else
          S.phase = (C.defl && S.dready && S.dden > 0.0) ? PH_DEFL_A : PH_FIN;
        break;
      }
      S.it = 0;
      S.inner_atol = pymax(C.inner_rtol * rn, 0.5 * S.atol);
      inner_start(C, S);
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
        S.phase = outer_entry(C);
      } else {
        S.phase = (sqrt(S.rnorm2) < S.inner_atol) ? outer_entry(C) : PH_CG;
      }
      break;
    case PH_CG: tail_cg(C, S, tot); break;
    case PH_DEFL_A: {
      const double cf = tot[0] / S.dden;
      S.dcoef = isfinite(cf) ? cf : 0.0;
      S.phase = PH_DEFL_B;
      break;
    }
    case PH_DEFL_B: S.phase = PH_FIN; break;
    case PH_DSTORE:  // the set-up's iterations and passes are not the step's
      S.dden = tot[0];
      S.dready = 1;
      S.it_run -= S.it_total;
      S.outer_run -= S.outer;
      enter_pinit(C, S);
      break;
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
      S.step_len = S.rounds_run - S.step_round;
      S.step_round = S.rounds_run;
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
  double dt, mean, beta, alpha, x0;  // x0: x[0] in FIN and CORR_A; the Galerkin coefficient in DEFL_B
  int it, x_is_b, par;  // par = round parity
  float cx, cy, betaf, alphaf, minvf;  // float32 path: 1/dx^2, 1/dy^2, CG scalars, Jacobi weight
  float ca0f, cb0f, ca1f, cb1f;        // two-iteration round: alpha_k, beta_k, alpha_k+1, beta_k+1
  float ca2f, cb2f;                    // three-iteration round: alpha_k+2, beta_k+2
  int ca_init;                         // two-iteration round that only forms the starting basis
  int dset;                            // PINIT / OUTER_B of the deflation set-up: the right-hand side is 1
  double ca0, ca1;                     // alpha_k, alpha_k+1 for an FP64 iterate
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
  double bb;
  if (A.dset) {  // deflation set-up: A y = 1 off the pinned cell
    bb = (k == 0) ? 0.0 : 1.0;
  } else {
    bb = C.b[k];
    if (C.mode != PF_MODE_SOLVE) {
      bb = bb - A.mean;
      if (k == 0) bb = 0.0;
      C.b[k] = bb;
    }
  }
  const T rr = (T)bb;
  ((T*)((C.prec == PF_PREC_MIXED && A.par) ? C.r1 : C.r))[k] = rr;  // MIXED: r[q] of round q
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
__device__ __attribute__((always_inline)) inline float minv_f32(const PfCav& C, const PhaseArgs& A, int j, int i) {
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
__device__ inline void cell_outer_b(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const size_t k = (size_t)j * C.nx + i;
  Grid<double> XG{C.x64, C.nx};
  const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
  const double bk = A.dset ? ((k == 0) ? 0.0 : 1.0) : C.b[k];
  const double r64 = bk - ax;
  const T rt = (T)r64;
  ((T*)(A.par ? C.r1 : C.r))[k] = rt;
  const T z = (T)jacobi_minv(C.jacobi, C.invd, C.nx, C.ny, j, i) * rt;
  acc[0] += r64 * r64;
  acc[1] += (double)rt * (double)z;
  acc[2] += (double)rt * (double)rt;
}

// Deflation: y.(b - A x) of the finished solve's FP64 iterate.
__device__ inline void cell_defl_a(const PfCav& C, int j, int i, double* acc) {
  const size_t k = (size_t)j * C.nx + i;
  Grid<double> XG{C.x64, C.nx};
  const double ax = apply_op_at<double>(C.op, XG, C.nx, C.ny, j, i, C.dx, C.dy);
  acc[0] += (double)C.y[k] * (C.b[k] - ax);
}

__device__ inline void cell_defl_b(const PfCav& C, const PhaseArgs& A, int j, int i) {
  const size_t k = (size_t)j * C.nx + i;
  C.x64[k] = C.x64[k] + A.x0 * (double)C.y[k];  // x0 carries the coefficient
}

// Deflation set-up: y = float(x) of the solve of A y = 1, and y.A y of the
// stored (rounded) y, with the stencil applied in FP64.
__device__ inline void cell_dstore(const PfCav& C, int j, int i, double* acc) {
  const int nx = C.nx;
  const size_t k = (size_t)j * nx + i;
  const double* X = C.x64;
  auto YG = [&](int jj, int ii) -> double { return (double)(float)X[(size_t)jj * nx + ii]; };
  const double ay = apply_op_at<double>(C.op, YG, nx, C.ny, j, i, C.dx, C.dy);
  const float yv = (float)X[k];
  C.y[k] = yv;
  acc[0] += (double)yv * ay;
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
  T* US = (T*)C.us;
  T* VS = (T*)C.vs;
  T uc, vc;
  if constexpr (sizeof(T) == sizeof(double)) {
    const T dt = (T)A.dt;
    const T gx = grad_x_coherent<T>(PG, nx, ny, j, i, (T)C.dx);
    const T gy = grad_y_coherent<T>(PG, nx, ny, j, i, (T)C.dy);
    uc = US[k] - dt * gx;
    vc = VS[k] - dt * gy;
  } else {
    // MIXED: the gradient of the FP64 solution p = x - x[0] (FIN stores p
    // rounded to float32; near the lid corners of an impulsive start p is
    // large and its float32 rounding would move u by dt * |p| * 2^-24).
    const double* XS = A.x_is_b ? C.b : C.x64;
    const double x0 = A.x0;
    auto P64 = [&](int jj, int ii) -> double { return XS[(size_t)jj * nx + ii] - x0; };
    const double gx = grad_x_coherent<double>(P64, nx, ny, j, i, C.dx);
    const double gy = grad_y_coherent<double>(P64, nx, ny, j, i, C.dy);
    uc = (T)((double)US[k] - A.dt * gx);
    vc = (T)((double)VS[k] - A.dt * gy);
  }
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

// MIXED inner PCG, single reduction per iteration (Chronopoulos & Gear, with
// s = A p applied explicitly to the stored search direction):
//   u_k = M r_k;  p_k = u_k + beta_k p_{k-1};  s_k = A p_k
//   x_{k+1} = x_k + alpha_k p_k;  r_{k+1} = r_k - alpha_k s_k
//   u_{k+1} = M r_{k+1};  w_{k+1} = A u_{k+1}
//   partials gamma = r_{k+1}.u_{k+1}, delta = w_{k+1}.u_{k+1}, |r_{k+1}|^2
// and the tail sets beta_{k+1} = gamma'/gamma, alpha_{k+1} = gamma'/(delta' -
// beta_{k+1} gamma'/alpha_k). Because the update of r uses A applied to the
// same rounded p that updates x, the recurred residual stays consistent with
// x up to the rounding of x itself. r and p are double-buffered by iteration
// parity; x is updated in place. CG0 computes the partials of k = 0.
template <class T>
__device__ inline void cell_cg0(const PfCav& C, const PhaseArgs& A, int j, int i, double* acc) {
  const int nx = C.nx, ny = C.ny;
  const T* R = (const T*)(A.par ? C.r : C.r1);  // r[q^1]: written by PINIT / OUTER_B in round q-1
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
  const T rr = R[(size_t)j * nx + i];
  acc[0] += (double)rr * (double)c;
  acc[1] += (double)wv * (double)c;
  acc[2] += (double)rr * (double)rr;
}

// One float32 PCG iteration over a workgroup's region (64 columns x RH = 16*M
// rows), staged through LDS with a 2-cell halo:
//   pass 1  (region_load) every global load of the thread, issued together:
//           r_k, p_{k-1}, x at its M cells and r_k, p_{k-1} at (at most) one
//           cell of the 2-cell ring;
//   pass 2  p_k into LDS (cells and ring), r_k of the ring into LDS; store
//           p_k and x_{k+1};
//   pass 3  s_k = A p_k, r_{k+1}, u_{k+1} at the cells and the 1-cell ring;
//           store r_{k+1}; u_{k+1} into LDS;
//   pass 4  w_{k+1} = A u_{k+1} and the partials at the cells.
// HBM traffic per cell: read r, p, x; write r, p, x (the halo re-reads hit
// L2). The r ring and u share one LDS array: a ring cell's r is read by the
// thread that then writes the same cell's u, and no other thread touches ring
// cells in pass 3.
// Buffers follow the round parity q: the iteration run in round q reads r[q],
// p[q^1] and writes r[q^1], p[q]; PINIT and OUTER_B (two rounds before the
// first iteration of a solve) write r[q] and CG0 reads r[q^1].
constexpr int HP = TX + 4;  // LDS row pitch: 64 columns + 2-cell halo each side
__host__ __device__ constexpr int region_lds_floats(int m) { return 2 * (TY * m + 4) * HP; }

__device__ inline float* mixed_r(const PfCav& C, int q) { return (float*)(q ? C.r1 : C.r); }
__device__ inline float* mixed_p(const PfCav& C, int q) { return (float*)(q ? C.d1 : C.d0); }

template <int M, bool XD>
struct RegionIn {
  float ro[M], po[M];
  std::conditional_t<XD, double, float> xo[M];  // iterate: x64 (XD) or the FP32 correction
  float hr, hp;
  int hrow, hcol;
};

template <int M, bool XD>
__device__ __attribute__((always_inline)) inline void region_load(const PfCav& C, int par, int r, RegionIn<M, XD>& L) {
  const int nx = C.nx, ny = C.ny;
  const float* __restrict__ R = mixed_r(C, par);
  const float* __restrict__ Pold = mixed_p(C, par ^ 1);
  const float* __restrict__ X = (const float*)C.x;
  constexpr int RH = TY * M;
  const int li = threadIdx.x, lj = threadIdx.y;
  const int tid = lj * TX + li;
  const int i0 = (r % C.tiles_x) * TX;
  const int j0 = (r / C.tiles_x) * RH;
  const int i = i0 + li;
  // Loads are unconditional (out-of-range lanes read cell 0 and discard it),
  // so none of them sits behind a branch and all are in flight together.
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    const bool ok = (i < nx) && (j < ny);
    const size_t k = ok ? (size_t)j * nx + i : 0;
    const float r = R[k], po = Pold[k];
    L.ro[m] = ok ? r : 0.f;
    L.po[m] = ok ? po : 0.f;
    if constexpr (XD) {
      const double x = C.x64[k];
      L.xo[m] = ok ? x : 0.0;
    } else {
      const float x = X[k];
      L.xo[m] = ok ? x : 0.f;
    }
  }
  // 2-cell ring: rows 0,1 and RH+2,RH+3 (all HP columns), then columns 0,1
  // and HP-2,HP-1 of rows 2..RH+1 (LDS coordinates row = j-j0+2, col = i-i0+2).
  int hrow = -1, hcol = 0;
  if (tid < 2 * HP) {
    hrow = tid / HP;
    hcol = tid % HP;
  } else if (tid < 4 * HP) {
    const int q = tid - 2 * HP;
    hrow = RH + 2 + q / HP;
    hcol = q % HP;
  } else if (tid < 4 * HP + 4 * RH) {
    const int q = tid - 4 * HP;
    const int side = q / (2 * RH), rem = q % (2 * RH);
    hrow = 2 + rem / 2;
    hcol = side ? TX + 2 + (rem & 1) : (rem & 1);
  }
  L.hrow = hrow;
  L.hcol = hcol;
  const int hj = j0 + hrow - 2, hi = i0 + hcol - 2;
  const bool hin = hrow >= 0 && hi >= 0 && hi < nx && hj >= 0 && hj < ny;
  const size_t kk = hin ? (size_t)hj * nx + hi : 0;
  const float hr = R[kk], hp = Pold[kk];
  L.hr = hin ? hr : 0.f;
  L.hp = hin ? hp : 0.f;
}

template <int M, bool XD>
__device__ __attribute__((always_inline)) inline void region_cg_f32(const PfCav& C, const PhaseArgs& A, int r, double* acc,
                                     float* __restrict__ lds, const RegionIn<M, XD>& L) {
  const int nx = C.nx, ny = C.ny;
  float* __restrict__ Rn = mixed_r(C, A.par ^ 1);
  float* __restrict__ Pn = mixed_p(C, A.par);
  float* __restrict__ X = (float*)C.x;
  const int first = (A.it == 0);
  const float al = A.alphaf, be = A.betaf;
  constexpr int RH = TY * M;
  constexpr int NR = RH + 4;
  float* __restrict__ lp = lds;            // p_k on the region + 2-cell ring
  float* __restrict__ lu = lds + NR * HP;  // r_k on the ring, then u_{k+1}
  const int li = threadIdx.x, lj = threadIdx.y;
  const int tid = lj * TX + li;
  const int i0 = (r % C.tiles_x) * TX;
  const int j0 = (r / C.tiles_x) * RH;
  const int i = i0 + li;
  // Pass 2.
  bool in[M];
  float pkr[M];  // p_k at the cells (FP64-iterate partials)
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    in[m] = (i < nx) && (j < ny);
    float pk = 0.f;
    if (in[m]) {
      const float u = minv_f32(C, A, j, i) * L.ro[m];
      pk = first ? u : (u + be * L.po[m]);
      const size_t k = (size_t)j * nx + i;
      // Streaming stores: these lines are next read in the next round, by
      // this and neighbouring workgroups, so keeping them dirty in L2 only
      // adds to the write-back at the end of the kernel.
      pcg_store(pk, &Pn[k]);
      if constexpr (XD)
        pcg_store(L.xo[m] + A.alpha * (double)pk, &C.x64[k]);
      else
        pcg_store(L.xo[m] + al * pk, &X[k]);
    }
    lp[(TY * m + lj + 2) * HP + li + 2] = pk;
    pkr[m] = pk;
  }
  if (L.hrow >= 0) {
    const int hj = j0 + L.hrow - 2, hi = i0 + L.hcol - 2;
    const bool hin = hi >= 0 && hi < nx && hj >= 0 && hj < ny;
    float p = 0.f;
    if (hin) {
      const float u = minv_f32(C, A, hj, hi) * L.hr;
      p = first ? u : (u + be * L.hp);
    }
    lp[L.hrow * HP + L.hcol] = p;
    lu[L.hrow * HP + L.hcol] = hin ? L.hr : 0.f;
  }
  __syncthreads();
  // Pass 3.
  auto s_at = [&](int row, int col, int j, int ii) -> float {
    const float* q = lp + row * HP + col;
    return apply_op5_fast(C.op, q[0], q[1], q[-1], q[HP], q[-HP], nx, ny, j, ii, A.cx, A.cy);
  };
  float rk1[M], uk1[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    const int row = TY * m + lj + 2;
    rk1[m] = uk1[m] = 0.f;
    if (in[m]) {
      const float sv = s_at(row, li + 2, j, i);
      rk1[m] = L.ro[m] - al * sv;
      uk1[m] = minv_f32(C, A, j, i) * rk1[m];
      if constexpr (XD) {
        acc[3] += (double)pkr[m] * (double)sv;
        acc[4] += (double)uk1[m] * (double)sv;
        acc[5] += (double)rk1[m] * (double)pkr[m];
      }
      pcg_store(rk1[m], &Rn[(size_t)j * nx + i]);
    }
    lu[row * HP + li + 2] = uk1[m];
  }
  // 1-cell ring: rows 1 and RH+2 (columns 1..TX+2), columns 1 and TX+2 (rows 2..RH+1)
  int grow = -1, gcol = 0;
  if (tid < 2 * (TX + 2)) {
    grow = (tid < TX + 2) ? 1 : RH + 2;
    gcol = 1 + tid % (TX + 2);
  } else if (tid < 2 * (TX + 2) + 2 * RH) {
    const int q = tid - 2 * (TX + 2);
    grow = 2 + q % RH;
    gcol = (q < RH) ? 1 : TX + 2;
  }
  if (grow >= 0) {
    const int gj = j0 + grow - 2, gi = i0 + gcol - 2;
    float u = 0.f;
    if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
      const float sv = s_at(grow, gcol, gj, gi);
      u = minv_f32(C, A, gj, gi) * (lu[grow * HP + gcol] - al * sv);
    }
    lu[grow * HP + gcol] = u;
  }
  __syncthreads();
  // Pass 4.
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (in[m]) {
      const int j = j0 + TY * m + lj;
      const float* q = lu + (TY * m + lj + 2) * HP + li + 2;
      const float wv = apply_op5_fast(C.op, q[0], q[1], q[-1], q[HP], q[-HP], nx, ny, j, i, A.cx, A.cy);
      acc[0] += (double)rk1[m] * (double)uk1[m];
      acc[1] += (double)wv * (double)uk1[m];
      acc[2] += (double)rk1[m] * (double)rk1[m];
    }
  }
}

// ---------------------------------------------------------------------------
// Two PCG iterations per round (cavities with PfCav::ca; the tail is tail_ca).
// A workgroup region of 64 x RH cells is staged in LDS with a 4-cell frame:
//   pass 1  loads r_k, p_k at the cells and the frame, x at the cells;
//           p (0 in the basis-only round) and z = M r into LDS;
//   pass 2  a = A p within distance 3; M a into LDS;
//   pass 3  c = A M a and e = A z within distance 2, then the two steps
//             p_k+1 = z - alpha_k M a + beta_k p,  A p_k+1 = e - alpha_k c + beta_k a
//             r_k+2 = r - alpha_k a - alpha_k+1 A p_k+1,  z_k+2 = M r_k+2
//             p_k+2 = z_k+2 + beta_k+1 p_k+1,  x_k+2 = x + alpha_k p + alpha_k+1 p_k+1;
//           stores r_k+2, p_k+2, x_k+2; p_k+2, z_k+2 into LDS;
//   pass 4  a' = A p_k+2 within distance 1; M a' into LDS;
//   pass 5  c' = A M a', e' = A z_k+2 and the Gram partials at the cells.
// The frame cells are recomputed from the same inputs with the same
// expressions as their owners compute them, so they carry the same values.
// HBM traffic per cell is that of the one-iteration round (read r, p, x;
// write r, p, x) for two iterations; the frame re-reads hit L2.
// LDS row pitch: 64 columns + 4-cell frame each side, plus one pad column.
// An odd pitch spreads a frame column (cells one pitch apart, one per lane)
// over all LDS banks; with 72 its cells fell on 4 banks. 13.78 -> 13.62 us
// per round at 1024^2.
constexpr int HQ = TX + 8 + 1;
__host__ __device__ constexpr int ca_lds_floats(int m) { return 5 * (TY * m + 8) * HQ; }
// Cells at Chebyshev distance exactly d from a region of rh rows, and within D.
__host__ __device__ constexpr int ca_ring_cells(int rh, int d) { return 2 * (TX + 2 * d) + 2 * (rh + 2 * d - 2); }
__host__ __device__ constexpr int ca_frame_cells(int rh, int D) {
  return D <= 0 ? 0 : ca_frame_cells(rh, D - 1) + ca_ring_cells(rh, D);
}

// LDS row/column of frame cell q, frame cells numbered ring by ring (distance
// 1, 2, 3, 4; each ring: top row, bottom row, left column, right column), so
// the cells within distance D are exactly q < ca_frame_cells(rh, D). Region
// cell (jj, ii) sits at row jj + 4, column ii + 4.
__host__ __device__ inline void ca_frame_cell(int rh, int q, int& row, int& col) {
  int d = 1;
  const int n1 = ca_ring_cells(rh, 1), n2 = ca_ring_cells(rh, 2), n3 = ca_ring_cells(rh, 3);
  if (q >= n1) {
    q -= n1;
    d = 2;
    if (q >= n2) {
      q -= n2;
      d = 3;
      if (q >= n3) {
        q -= n3;
        d = 4;
      }
    }
  }
  const int w = TX + 2 * d, hs = rh + 2 * d - 2;
  if (q < w) {
    row = 4 - d;
    col = 4 - d + q;
  } else if (q < 2 * w) {
    row = 3 + rh + d;
    col = 4 - d + (q - w);
  } else if (q < 2 * w + hs) {
    col = 4 - d;
    row = 5 - d + (q - 2 * w);
  } else {
    col = 3 + TX + d;
    row = 5 - d + (q - 2 * w - hs);
  }
}

// The two steps at one cell (pass 3), shared by cells and frame.
struct CaCell {
  float p1, r2, z2, p2;
};
__device__ __attribute__((always_inline)) inline CaCell ca_steps(const PhaseArgs& A, float m, float r, float p, float a,
                                                                 float c, float e) {
  CaCell o;
  const float z = m * r;
  const float ma = m * a;
  o.p1 = (z - A.ca0f * ma) + A.cb0f * p;
  const float ap1 = (e - A.ca0f * c) + A.cb0f * a;
  o.r2 = (r - A.ca0f * a) - A.ca1f * ap1;
  o.z2 = m * o.r2;
  o.p2 = o.z2 + A.cb1f * o.p1;
  return o;
}

// XD: the iterate is the FP64 x64 (x64 += alpha_k p_k + alpha_k+1 p_k+1 in
// FP64), as in the one-iteration FP64-iterate round.
template <int M, bool XD>
__device__ __attribute__((always_inline)) inline void region_ca(const PfCav& C, const PhaseArgs& A, int r, double* acc,
                                                                float* __restrict__ lds) {
  const int nx = C.nx, ny = C.ny;
  constexpr int RH = TY * M;
  constexpr int NR = RH + 8;
  constexpr int F1 = ca_frame_cells(RH, 1), F2 = ca_frame_cells(RH, 2), F3 = ca_frame_cells(RH, 3);
  constexpr int F4 = ca_frame_cells(RH, 4);
  constexpr int NS = (F4 + NTH - 1) / NTH;  // frame slots per thread; slots past the first are loads only
  static_assert(F3 <= NTH, "one thread per frame cell within distance 3");
  static_assert(NR * HQ * 5 == ca_lds_floats(M), "LDS layout");
  float* __restrict__ lP = lds;                 // p_k
  float* __restrict__ lZ = lds + NR * HQ;       // z_k = M r_k
  float* __restrict__ lMA = lds + 2 * NR * HQ;  // M A p_k, then M A p_k+2
  float* __restrict__ lP2 = lds + 3 * NR * HQ;  // p_k+2
  float* __restrict__ lZ2 = lds + 4 * NR * HQ;  // z_k+2
  // Buffers follow the round parity q as in region_cg_f32: read r[q], p[q^1],
  // write r[q^1], p[q]. The basis-only round runs the round after PINIT /
  // OUTER_B, which wrote r[q^1]: it reads r there, leaves r and x as they
  // are (the next round reads r[q^1] again) and writes p = M r to p[q].
  const bool init = A.ca_init != 0;
  const float* __restrict__ R = mixed_r(C, init ? (A.par ^ 1) : A.par);
  const float* __restrict__ Pold = mixed_p(C, A.par ^ 1);
  float* __restrict__ Rn = mixed_r(C, A.par ^ 1);
  float* __restrict__ Pn = mixed_p(C, A.par);
  float* __restrict__ X = (float*)C.x;
  const int li = threadIdx.x, lj = threadIdx.y;
  const int tid = lj * TX + li;
  const int i0 = (r % C.tiles_x) * TX;
  const int j0 = (r / C.tiles_x) * RH;
  const int i = i0 + li;
  // Operator without boundary tests. A cell on the domain boundary also
  // writes its value into the LDS ghost across that boundary, so the missing
  // neighbour's difference (e - c etc.) is exactly +0, as the reflective
  // stencil drops it (lap5_fast); the operator's sign is folded into
  // 1/dx^2, 1/dy^2 (round-to-nearest is sign-symmetric). Identity rows (the
  // pinned cell; the Dirichlet boundary) return the cell's value. Cells
  // outside the domain are never written and never read.
  const bool pinned = (C.op == 0);
  const float scx = pinned ? -A.cx : A.cx, scy = pinned ? -A.cy : A.cy;
  // Cell code: LDS offsets of the (at most two) ghosts the cell writes, the
  // cell's own offset standing in for a missing one; bit 30 marks a cell with
  // a ghost (only those write the copies: 4096^2 207 instead of 217 us per
  // round, 1024^2 13.7 instead of 14.2 us), bit 31 an identity row.
  auto code_of = [&](int gj, int gi, int row, int col) -> unsigned {
    const int o = row * HQ + col;
    int g1 = o, g2 = o;
    if (gi == nx - 1 && col + 1 < HQ) g1 = o + 1;
    if (gi == 0 && col >= 1) g1 = o - 1;
    if (gj == ny - 1 && row + 1 < NR) g2 = o + HQ;
    if (gj == 0 && row >= 1) g2 = o - HQ;
    const bool id = pinned ? (gj == 0 && gi == 0) : (gj == 0 || gj == ny - 1 || gi == 0 || gi == nx - 1);
    return (unsigned)g1 | ((unsigned)g2 << 15) | (id ? 0x80000000u : 0u) | ((g1 != o || g2 != o) ? 0x40000000u : 0u);
  };
  static_assert(NR * HQ < (1 << 15), "LDS offsets fit 15 bits");
  auto put = [&](float* L, int o, unsigned code, float v) {
    L[o] = v;
    if (code & 0x40000000u) {
      L[code & 0x7fffu] = v;
      L[(code >> 15) & 0x7fffu] = v;
    }
  };
  auto opv = [&](const float* L, int o, unsigned code) -> float {
    const float c = L[o];
    const float t = __fmul_rn((L[o + 1] - c) - (c - L[o - 1]), scx) + __fmul_rn((L[o + HQ] - c) - (c - L[o - HQ]), scy);
    return (code & 0x80000000u) ? c : t;
  };
  // Pass 1: every global load of the thread, unconditional (out-of-range
  // lanes read cell 0 and discard it).
  using XT = std::conditional_t<XD, double, float>;
  float ro[M], po[M];
  XT xo[M];
  bool in[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    in[m] = (i < nx) && (j < ny);
    const size_t k = in[m] ? (size_t)j * nx + i : 0;
    const float rv = R[k], pv = Pold[k];
    XT xv;
    if constexpr (XD)
      xv = C.x64[k];
    else
      xv = X[k];
    ro[m] = in[m] ? rv : 0.f;
    po[m] = (in[m] && !init) ? pv : 0.f;
    xo[m] = in[m] ? xv : XT(0);
  }
  float fr[NS], fp[NS];
  int fo[NS];
  unsigned fc[NS];
  bool fin[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int q = tid + s * NTH;
    int row = 0, col = 0;
    if (q < F4) ca_frame_cell(RH, q, row, col);
    fo[s] = row * HQ + col;
    const int gj = j0 + row - 4, gi = i0 + col - 4;
    fin[s] = (q < F4) && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
    fc[s] = fin[s] ? code_of(gj, gi, row, col) : 0;
    const size_t kk = fin[s] ? (size_t)gj * nx + gi : 0;
    const float rv = R[kk], pv = Pold[kk];
    fr[s] = fin[s] ? rv : 0.f;
    fp[s] = (fin[s] && !init) ? pv : 0.f;
  }
  const int o0 = (lj + 4) * HQ + li + 4;  // LDS offset of the thread's first cell; cell m at o0 + m*TY*HQ
  float mo[M];     // Jacobi weights of the cells
  unsigned cc[M];  // ghost / identity codes of the cells
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    mo[m] = in[m] ? minv_f32(C, A, j, i) : 0.f;
    cc[m] = in[m] ? code_of(j, i, TY * m + lj + 4, li + 4) : 0;
    if (in[m]) {
      put(lP, o0 + m * TY * HQ, cc[m], po[m]);
      put(lZ, o0 + m * TY * HQ, cc[m], mo[m] * ro[m]);
    }
  }
  float fm0 = 0.f;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (fin[s]) {
      const int row = fo[s] / HQ, col = fo[s] - row * HQ;
      const float fm = minv_f32(C, A, j0 + row - 4, i0 + col - 4);
      if (s == 0) fm0 = fm;
      put(lP, fo[s], fc[s], fp[s]);
      put(lZ, fo[s], fc[s], fm * fr[s]);
    }
  }
  __syncthreads();
  // Pass 2: a = A p.
  float ao[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    ao[m] = 0.f;
    if (in[m]) {
      ao[m] = opv(lP, o0 + m * TY * HQ, cc[m]);
      put(lMA, o0 + m * TY * HQ, cc[m], mo[m] * ao[m]);
    }
  }
  float fa = 0.f;
  if (tid < F3 && fin[0]) {
    fa = opv(lP, fo[0], fc[0]);
    put(lMA, fo[0], fc[0], fm0 * fa);
  }
  __syncthreads();
  // Pass 3: the two steps.
  float r2o[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    r2o[m] = 0.f;
    if (in[m]) {
      const int o = o0 + m * TY * HQ;
      const float cv = opv(lMA, o, cc[m]);
      const float ev = opv(lZ, o, cc[m]);
      const CaCell v = ca_steps(A, mo[m], ro[m], po[m], ao[m], cv, ev);
      const size_t k = (size_t)(j0 + TY * m + lj) * nx + i;
      // Streaming stores: the lines are next read in the next round.
      pcg_store(v.p2, &Pn[k]);
      if (!init) {
        pcg_store(v.r2, &Rn[k]);
        if constexpr (XD)
          pcg_store((xo[m] + A.ca0 * (double)po[m]) + A.ca1 * (double)v.p1, &C.x64[k]);
        else
          pcg_store((xo[m] + A.ca0f * po[m]) + A.ca1f * v.p1, &X[k]);
      }
      r2o[m] = v.r2;
      put(lP2, o, cc[m], v.p2);
      put(lZ2, o, cc[m], v.z2);
    }
  }
  if (tid < F2 && fin[0]) {
    const float cv = opv(lMA, fo[0], fc[0]);
    const float ev = opv(lZ, fo[0], fc[0]);
    const CaCell v = ca_steps(A, fm0, fr[0], fp[0], fa, cv, ev);
    put(lP2, fo[0], fc[0], v.p2);
    put(lZ2, fo[0], fc[0], v.z2);
  }
  __syncthreads();
  // Pass 4: a' = A p_k+2 (lMA is rewritten: pass 3's reads of it are done).
  float a2o[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    a2o[m] = 0.f;
    if (in[m]) {
      a2o[m] = opv(lP2, o0 + m * TY * HQ, cc[m]);
      put(lMA, o0 + m * TY * HQ, cc[m], mo[m] * a2o[m]);
    }
  }
  if (tid < F1 && fin[0]) put(lMA, fo[0], fc[0], fm0 * opv(lP2, fo[0], fc[0]));
  __syncthreads();
  // Pass 5: partials of the new basis at the cells. Products of two floats
  // are exact in FP64, so each fma equals the product followed by the sum.
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (in[m]) {
      const int o = o0 + m * TY * HQ;
      const float cv = opv(lMA, o, cc[m]);
      const float ev = opv(lZ2, o, cc[m]);
      const double p2 = lP2[o], z2 = lZ2[o], r2 = r2o[m], a2 = a2o[m];
      const double ma2 = mo[m] * a2o[m], c2 = cv, mc2 = mo[m] * cv, e2 = ev, me2 = mo[m] * ev;
      acc[0] = __fma_rn(p2, a2, acc[0]);     // W01
      acc[1] = __fma_rn(p2, c2, acc[1]);     // W02
      acc[2] = __fma_rn(p2, e2, acc[2]);     // W04
      acc[3] = __fma_rn(a2, ma2, acc[3]);    // W11
      acc[4] = __fma_rn(a2, mc2, acc[4]);    // W12
      acc[5] = __fma_rn(a2, z2, acc[5]);     // W13
      acc[6] = __fma_rn(a2, me2, acc[6]);    // W14
      acc[7] = __fma_rn(c2, mc2, acc[7]);    // W22
      acc[8] = __fma_rn(c2, z2, acc[8]);     // W23
      acc[9] = __fma_rn(c2, me2, acc[9]);    // W24
      acc[10] = __fma_rn(r2, z2, acc[10]);   // W33
      acc[11] = __fma_rn(z2, e2, acc[11]);   // W34
      acc[12] = __fma_rn(e2, me2, acc[12]);  // W44
      acc[13] = __fma_rn(r2, r2, acc[13]);   // |r|^2
    }
  }
}

// Three PCG iterations per round (cavities with PfCav::ca == 2; the tail is
// tail_ca3), s-step PCG with s = 3 on the basis
// Y = [p, Bp, B^2 p, B^3 p, z, Bz, B^2 z], B = M A. Regions carry a 6-cell
// frame; four LDS buffers L0..L3 rotate through the passes:
//   pass 1  loads r_k, p_k at the cells and the frame (distance <= 6), x at
//           the cells; p -> L0, z = M r -> L1;
//   pass 2  a = A p (distance <= 5), e = A z (<= 4); M a -> L2, M e -> L3;
//   pass 3  c = A M a (<= 4), f = A M e (<= 3); M c -> L0;
//   pass 4  g = A M c (<= 3) and the three steps (<= 3)
//             p1 = (z - a0 M a) + b0 p,          A p1 = (e - a0 c) + b0 a,
//             r2 = (r - a0 a) - a1 A p1,  z2 = M r2,  p2 = z2 + b1 p1,
//             A M A p1 = (f - a0 g) + b0 c,
//             A p2 = ((e - a0 c) - a1 A M A p1) + b1 A p1,
//             r3 = r2 - a2 A p2,  z3 = M r3,  p3 = z3 + b2 p2,
//             x3 = ((x + a0 p) + a1 p1) + a2 p2;
//           stores r3, p3, x3; p3 -> L2, z3 -> L3;
//   pass 5  a' = A p3 (<= 2), e' = A z3 (<= 1); M a' -> L0, M e' -> L1;
//   pass 6  c' = A M a' (<= 1), f' = A M e' (cells); M c' -> L2;
//   pass 7  g' = A M c' and, at the cells, the 26 entries of W that the
//           tail reads plus |r3|^2 as FP64 partials.
// As in the two-iteration round, frame cells are computed from the same
// inputs with the same expressions as their owners, and each of r, p, x is
// read and written once per round, now for three iterations.
constexpr int FW3 = 6;              // frame width of the three-iteration round
constexpr int HQ3 = TX + 2 * FW3 + 1;  // LDS row pitch, odd as HQ
__host__ __device__ constexpr int ca3_lds_floats(int m) { return 4 * (TY * m + 2 * FW3) * HQ3; }

// LDS row/column of frame cell q of a region with a frame of FW cells,
// numbered ring by ring as ca_frame_cell does; region cell (jj, ii) sits at
// row jj + FW, column ii + FW.
template <int FW>
__host__ __device__ inline void frame_cell_w(int rh, int q, int& row, int& col) {
  int d = 1;
  for (; d < FW; ++d) {
    const int nd = ca_ring_cells(rh, d);
    if (q < nd) break;
    q -= nd;
  }
  const int w = TX + 2 * d, hs = rh + 2 * d - 2;
  if (q < w) {
    row = FW - d;
    col = FW - d + q;
  } else if (q < 2 * w) {
    row = FW - 1 + rh + d;
    col = FW - d + (q - w);
  } else if (q < 2 * w + hs) {
    col = FW - d;
    row = FW + 1 - d + (q - 2 * w);
  } else {
    col = FW - 1 + TX + d;
    row = FW + 1 - d + (q - 2 * w - hs);
  }
}

// The three steps at one cell (pass 4), shared by cells and frame.
struct Ca3Cell {
  float p1, p2, r3, z3, p3;
};
__device__ __attribute__((always_inline)) inline Ca3Cell ca3_steps(const PhaseArgs& A, float m, float r, float p,
                                                                   float a, float e, float c, float f, float g) {
  Ca3Cell o;
  const float z = m * r;
  const float ma = m * a;
  o.p1 = (z - A.ca0f * ma) + A.cb0f * p;
  const float ap1 = (e - A.ca0f * c) + A.cb0f * a;
  const float r2 = (r - A.ca0f * a) - A.ca1f * ap1;
  const float z2 = m * r2;
  o.p2 = z2 + A.cb1f * o.p1;
  const float amap1 = (f - A.ca0f * g) + A.cb0f * c;
  const float az1 = e - A.ca0f * c;
  const float ap2 = (az1 - A.ca1f * amap1) + A.cb1f * ap1;
  o.r3 = r2 - A.ca2f * ap2;
  o.z3 = m * o.r3;
  o.p3 = o.z3 + A.cb2f * o.p2;
  return o;
}

template <int M>
__device__ __attribute__((always_inline)) inline void region_ca3(const PfCav& C, const PhaseArgs& A, int r, double* acc,
                                                                 float* __restrict__ lds) {
  const int nx = C.nx, ny = C.ny;
  constexpr int RH = TY * M;
  constexpr int NR = RH + 2 * FW3;
  constexpr int F1 = ca_frame_cells(RH, 1), F2 = ca_frame_cells(RH, 2), F3 = ca_frame_cells(RH, 3);
  constexpr int F4 = ca_frame_cells(RH, 4), F5 = ca_frame_cells(RH, 5), F6 = ca_frame_cells(RH, 6);
  constexpr int NS = (F6 + NTH - 1) / NTH;  // frame slots per thread
  static_assert(F3 <= NTH && NS == 2, "slot 0 covers distance 3; two slots cover the frame");
  static_assert(NR * HQ3 * 4 == ca3_lds_floats(M), "LDS layout");
  static_assert(NR * HQ3 < (1 << 15), "LDS offsets fit 15 bits");
  float* __restrict__ L0 = lds;
  float* __restrict__ L1 = lds + NR * HQ3;
  float* __restrict__ L2 = lds + 2 * NR * HQ3;
  float* __restrict__ L3 = lds + 3 * NR * HQ3;
  // Buffers follow the round parity q as in region_ca: read r[q], p[q^1],
  // write r[q^1], p[q]; the basis-only round reads r[q^1], writes only p[q].
  const bool init = A.ca_init != 0;
  const float* __restrict__ R = mixed_r(C, init ? (A.par ^ 1) : A.par);
  const float* __restrict__ Pold = mixed_p(C, A.par ^ 1);
  float* __restrict__ Rn = mixed_r(C, A.par ^ 1);
  float* __restrict__ Pn = mixed_p(C, A.par);
  float* __restrict__ X = (float*)C.x;
  const int li = threadIdx.x, lj = threadIdx.y;
  const int tid = lj * TX + li;
  const int i0 = (r % C.tiles_x) * TX;
  const int j0 = (r / C.tiles_x) * RH;
  const int i = i0 + li;
  // Operator without boundary tests, as in region_ca: boundary cells mirror
  // their value into the LDS ghost across the boundary; bit 30 of a code
  // marks a cell with ghosts, bit 31 an identity row.
  const bool pinned = (C.op == 0);
  const float scx = pinned ? -A.cx : A.cx, scy = pinned ? -A.cy : A.cy;
  auto code_of = [&](int gj, int gi, int row, int col) -> unsigned {
    const int o = row * HQ3 + col;
    int g1 = o, g2 = o;
    if (gi == nx - 1 && col + 1 < HQ3) g1 = o + 1;
    if (gi == 0 && col >= 1) g1 = o - 1;
    if (gj == ny - 1 && row + 1 < NR) g2 = o + HQ3;
    if (gj == 0 && row >= 1) g2 = o - HQ3;
    const bool id = pinned ? (gj == 0 && gi == 0) : (gj == 0 || gj == ny - 1 || gi == 0 || gi == nx - 1);
    return (unsigned)g1 | ((unsigned)g2 << 15) | (id ? 0x80000000u : 0u) | ((g1 != o || g2 != o) ? 0x40000000u : 0u);
  };
  auto put = [&](float* L, int o, unsigned code, float v) {
    L[o] = v;
    if (code & 0x40000000u) {
      L[code & 0x7fffu] = v;
      L[(code >> 15) & 0x7fffu] = v;
    }
  };
  auto opv = [&](const float* L, int o, unsigned code) -> float {
    const float c = L[o];
    const float t =
        __fmul_rn((L[o + 1] - c) - (c - L[o - 1]), scx) + __fmul_rn((L[o + HQ3] - c) - (c - L[o - HQ3]), scy);
    return (code & 0x80000000u) ? c : t;
  };
  // Pass 1: every global load of the thread, unconditional (out-of-range
  // lanes read cell 0 and discard it).
  float ro[M], po[M], xo[M];
  bool in[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    in[m] = (i < nx) && (j < ny);
    const size_t k = in[m] ? (size_t)j * nx + i : 0;
    const float rv = R[k], pv = Pold[k], xv = X[k];
    ro[m] = in[m] ? rv : 0.f;
    po[m] = (in[m] && !init) ? pv : 0.f;
    xo[m] = in[m] ? xv : 0.f;
  }
  float fr[NS], fp[NS], fm[NS];
  int fo[NS];
  unsigned fc[NS];
  bool fin[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int q = tid + s * NTH;
    int row = 0, col = 0;
    if (q < F6) frame_cell_w<FW3>(RH, q, row, col);
    fo[s] = row * HQ3 + col;
    const int gj = j0 + row - FW3, gi = i0 + col - FW3;
    fin[s] = (q < F6) && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
    fc[s] = fin[s] ? code_of(gj, gi, row, col) : 0;
    const size_t kk = fin[s] ? (size_t)gj * nx + gi : 0;
    const float rv = R[kk], pv = Pold[kk];
    fr[s] = fin[s] ? rv : 0.f;
    fp[s] = (fin[s] && !init) ? pv : 0.f;
  }
  const int o0 = (lj + FW3) * HQ3 + li + FW3;  // LDS offset of the thread's first cell; cell m at o0 + m*TY*HQ3
  float mo[M];
  unsigned cc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = j0 + TY * m + lj;
    mo[m] = in[m] ? minv_f32(C, A, j, i) : 0.f;
    cc[m] = in[m] ? code_of(j, i, TY * m + lj + FW3, li + FW3) : 0;
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    fm[s] = 0.f;
    if (fin[s]) {
      const int row = fo[s] / HQ3, col = fo[s] - row * HQ3;
      fm[s] = minv_f32(C, A, j0 + row - FW3, i0 + col - FW3);
    }
  }
  // Per cell after pass 4: r3, p3, z3; after passes 5, 6: a', e', c', f'.
  float r3o[M], p3o[M], z3o[M];
  if (!init) {
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (in[m]) {
        put(L0, o0 + m * TY * HQ3, cc[m], po[m]);
        put(L1, o0 + m * TY * HQ3, cc[m], mo[m] * ro[m]);
      }
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (fin[s]) {
        put(L0, fo[s], fc[s], fp[s]);
        put(L1, fo[s], fc[s], fm[s] * fr[s]);
      }
    __syncthreads();
    // Pass 2: a = A p, e = A z.
    float ao[M], eo[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      ao[m] = eo[m] = 0.f;
      if (in[m]) {
        const int o = o0 + m * TY * HQ3;
        ao[m] = opv(L0, o, cc[m]);
        eo[m] = opv(L1, o, cc[m]);
        put(L2, o, cc[m], mo[m] * ao[m]);
        put(L3, o, cc[m], mo[m] * eo[m]);
      }
    }
    float fa = 0.f, fe = 0.f;  // slot 0 (distance <= 3 needs them in pass 4)
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int q = tid + s * NTH;
      if (fin[s] && q < F5) {
        const float av = opv(L0, fo[s], fc[s]);
        put(L2, fo[s], fc[s], fm[s] * av);
        if (s == 0) fa = av;
      }
      if (fin[s] && q < F4) {
        const float ev = opv(L1, fo[s], fc[s]);
        put(L3, fo[s], fc[s], fm[s] * ev);
        if (s == 0) fe = ev;
      }
    }
    __syncthreads();
    // Pass 3: c = A M a, f = A M e.
    float co[M], fo_[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      co[m] = fo_[m] = 0.f;
      if (in[m]) {
        const int o = o0 + m * TY * HQ3;
        co[m] = opv(L2, o, cc[m]);
        fo_[m] = opv(L3, o, cc[m]);
        put(L0, o, cc[m], mo[m] * co[m]);
      }
    }
    float fcv = 0.f, ffv = 0.f;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int q = tid + s * NTH;
      if (fin[s] && q < F4) {
        const float cv = opv(L2, fo[s], fc[s]);
        put(L0, fo[s], fc[s], fm[s] * cv);
        if (s == 0) fcv = cv;
      }
    }
    if (fin[0] && tid < F3) ffv = opv(L3, fo[0], fc[0]);
    __syncthreads();
    // Pass 4: g = A M c and the three steps.
#pragma unroll
    for (int m = 0; m < M; ++m) {
      r3o[m] = p3o[m] = z3o[m] = 0.f;
      if (in[m]) {
        const int o = o0 + m * TY * HQ3;
        const float gv = opv(L0, o, cc[m]);
        const Ca3Cell v = ca3_steps(A, mo[m], ro[m], po[m], ao[m], eo[m], co[m], fo_[m], gv);
        const size_t k = (size_t)(j0 + TY * m + lj) * nx + i;
        // Streaming stores: the lines are next read in the next round.
        pcg_store(v.p3, &Pn[k]);
        pcg_store(v.r3, &Rn[k]);
        pcg_store(((xo[m] + A.ca0f * po[m]) + A.ca1f * v.p1) + A.ca2f * v.p2, &X[k]);
        r3o[m] = v.r3;
        p3o[m] = v.p3;
        z3o[m] = v.z3;
        put(L2, o, cc[m], v.p3);
        put(L3, o, cc[m], v.z3);
      }
    }
    if (fin[0] && tid < F3) {
      const float gv = opv(L0, fo[0], fc[0]);
      const Ca3Cell v = ca3_steps(A, fm[0], fr[0], fp[0], fa, fe, fcv, ffv, gv);
      put(L2, fo[0], fc[0], v.p3);
      put(L3, fo[0], fc[0], v.z3);
    }
  } else {
    // Basis-only round: p3 = z3 = M r, r3 = r; p is stored, r and x stay.
#pragma unroll
    for (int m = 0; m < M; ++m) {
      r3o[m] = p3o[m] = z3o[m] = 0.f;
      if (in[m]) {
        const int o = o0 + m * TY * HQ3;
        const float zv = mo[m] * ro[m];
        const size_t k = (size_t)(j0 + TY * m + lj) * nx + i;
        pcg_store(zv, &Pn[k]);
        r3o[m] = ro[m];
        p3o[m] = zv;
        z3o[m] = zv;
        put(L2, o, cc[m], zv);
        put(L3, o, cc[m], zv);
      }
    }
    if (fin[0] && tid < F3) {
      const float zv = fm[0] * fr[0];
      put(L2, fo[0], fc[0], zv);
      put(L3, fo[0], fc[0], zv);
    }
  }
  __syncthreads();
  // Pass 5: a' = A p3, e' = A z3.
  float a2o[M], e2o[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    a2o[m] = e2o[m] = 0.f;
    if (in[m]) {
      const int o = o0 + m * TY * HQ3;
      a2o[m] = opv(L2, o, cc[m]);
      e2o[m] = opv(L3, o, cc[m]);
      put(L0, o, cc[m], mo[m] * a2o[m]);
      put(L1, o, cc[m], mo[m] * e2o[m]);
    }
  }
  if (fin[0] && tid < F2) put(L0, fo[0], fc[0], fm[0] * opv(L2, fo[0], fc[0]));
  if (fin[0] && tid < F1) put(L1, fo[0], fc[0], fm[0] * opv(L3, fo[0], fc[0]));
  __syncthreads();
  // Pass 6: c' = A M a', f' = A M e'.
  float c2o[M], f2o[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    c2o[m] = f2o[m] = 0.f;
    if (in[m]) {
      const int o = o0 + m * TY * HQ3;
      c2o[m] = opv(L0, o, cc[m]);
      f2o[m] = opv(L1, o, cc[m]);
      put(L2, o, cc[m], mo[m] * c2o[m]);
    }
  }
  if (fin[0] && tid < F1) put(L2, fo[0], fc[0], fm[0] * opv(L0, fo[0], fc[0]));
  __syncthreads();
  // Pass 7: g' = A M c' and the partials at the cells. y = [p3, M a', M c',
  // M g', z3, M e', M f'] and u = M^-1 y = [., a', c', g', r3, e', f']; the
  // entry W_ij is y_i.u_j, with z3 standing in for M r3 beside a', c', g'.
  // Products of two floats are exact in FP64, so each fma equals the
  // product followed by the sum.
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (in[m]) {
      const int o = o0 + m * TY * HQ3;
      const float gv = opv(L2, o, cc[m]);
      const float mv = mo[m];
      const double p3 = p3o[m], z3 = z3o[m], r3 = r3o[m];
      const double a = a2o[m], c = c2o[m], g = gv, e = e2o[m], f = f2o[m];
      const double ma = mv * a2o[m], mc = mv * c2o[m], mg = mv * gv, me = mv * e2o[m], mf = mv * f2o[m];
      acc[0] = __fma_rn(p3, a, acc[0]);     // W01
      acc[1] = __fma_rn(p3, c, acc[1]);     // W02
      acc[2] = __fma_rn(p3, g, acc[2]);     // W03
      acc[3] = __fma_rn(p3, e, acc[3]);     // W05
      acc[4] = __fma_rn(p3, f, acc[4]);     // W06
      acc[5] = __fma_rn(ma, a, acc[5]);     // W11
      acc[6] = __fma_rn(ma, c, acc[6]);     // W12
      acc[7] = __fma_rn(ma, g, acc[7]);     // W13
      acc[8] = __fma_rn(a, z3, acc[8]);     // W14
      acc[9] = __fma_rn(ma, e, acc[9]);     // W15
      acc[10] = __fma_rn(ma, f, acc[10]);   // W16
      acc[11] = __fma_rn(mc, c, acc[11]);   // W22
      acc[12] = __fma_rn(mc, g, acc[12]);   // W23
      acc[13] = __fma_rn(c, z3, acc[13]);   // W24
      acc[14] = __fma_rn(mc, e, acc[14]);   // W25
      acc[15] = __fma_rn(mc, f, acc[15]);   // W26
      acc[16] = __fma_rn(mg, g, acc[16]);   // W33
      acc[17] = __fma_rn(g, z3, acc[17]);   // W34
      acc[18] = __fma_rn(mg, e, acc[18]);   // W35
      acc[19] = __fma_rn(mg, f, acc[19]);   // W36
      acc[20] = __fma_rn(r3, z3, acc[20]);  // W44
      acc[21] = __fma_rn(z3, e, acc[21]);   // W45
      acc[22] = __fma_rn(z3, f, acc[22]);   // W46
      acc[23] = __fma_rn(me, e, acc[23]);   // W55
      acc[24] = __fma_rn(me, f, acc[24]);   // W56
      acc[25] = __fma_rn(mf, f, acc[25]);   // W66
      acc[26] = __fma_rn(r3, r3, acc[26]);  // |r3|^2
    }
  }
}

// Workgroup partials of the two-iteration round (KCA values per thread): a
// transposed wave reduction. Each exchange pairs a lane with one whose index
// differs in one more bit (lane distance 32 and 16 by the gfx950 permlane
// swaps, then the DPP row mirror (index xor 15) and half-row mirror (xor 7)); the
// lane keeps half of the values it carries and adds the partner's copy of
// that half, so 16 values become 8, 4, 2, 1, and the quad sums (xor 2, xor 1)
// finish. 17 FP64 additions per lane instead of 6 per value; lane 4q ends with
// value q. Then thread k < KCA folds the NWAVE wave totals in wave order.
template <int CTRL>
__device__ __attribute__((always_inline)) inline double dpp_full(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
template <bool HALVES>
__device__ __attribute__((always_inline)) inline double swap_sum(double x, double y) {
  const unsigned long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
  const unsigned xl = (unsigned)xb, xh = (unsigned)(xb >> 32), yl = (unsigned)yb, yh = (unsigned)(yb >> 32);
  const auto lo = HALVES ? __builtin_amdgcn_permlane32_swap(xl, yl, false, false)
                         : __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
  const auto hi = HALVES ? __builtin_amdgcn_permlane32_swap(xh, yh, false, false)
                         : __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
  const double xs = __longlong_as_double((long long)((unsigned long long)lo[0] | ((unsigned long long)hi[0] << 32)));
  const double ys = __longlong_as_double((long long)((unsigned long long)lo[1] | ((unsigned long long)hi[1] << 32)));
  return xs + ys;
}
__device__ inline void block_partials_ca(double (&v)[NP], double (*sm)[NPS], double* dst, int stride) {
  static_assert(KCA <= 16 && NP >= 16, "16 slots carried through the exchange");
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = KCA; k < 16; ++k) v[k] = 0.0;
  double a[8], b[4], c[2];
  // Distances 32 and 16: v_permlane32_swap / v_permlane16_swap exchange the
  // upper half (odd rows) of x with the lower half (even rows) of y, so each
  // lane ends with its own kept value and its partner's copy of it, and the
  // sum needs no select (FP64 addition is commutative).
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = swap_sum<true>(v[k], v[8 + k]);
#pragma unroll
  for (int k = 0; k < 4; ++k) b[k] = swap_sum<false>(a[k], a[4 + k]);
  {
    const bool hi = lane & 8;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double keep = hi ? b[2 + k] : b[k];
      const double give = hi ? b[k] : b[2 + k];
      c[k] = keep + dpp_full<0x140>(give);  // row_mirror
    }
  }
  double d;
  {
    const bool hi = lane & 4;
    const double keep = hi ? c[1] : c[0];
    const double give = hi ? c[0] : c[1];
    d = keep + dpp_full<0x141>(give);  // row_half_mirror
  }
  d = d + dpp_full<0x4E>(d);  // quad_perm [2,3,0,1]
  d = d + dpp_full<0xB1>(d);  // quad_perm [1,0,3,2]
  const int q = lane >> 2;
  if ((lane & 3) == 0 && q < KCA) sm[wave][q] = d;
  __syncthreads();
  if (tid < KCA) {
    double t = sm[0][tid];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) t += sm[w][tid];
    dst[(size_t)tid * stride] = t;
  }
}

// Workgroup partials of the three-iteration round: the transposed wave
// reduction of block_partials_ca with 32 slots (27 used): exchanges at lane
// distance 32 and 16 (permlane swaps), the row mirror, the half-row mirror
// and quad_perm [2,3,0,1] halve the values carried, quad_perm [1,0,3,2]
// finishes; lane 2q ends with value q.
__device__ inline void block_partials_ca3(double (&v)[32], double (*sm)[NPS], double* dst, int stride) {
  static_assert(KCA3 <= 32 && NPS >= 32, "32 slots carried through the exchange");
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = KCA3; k < 32; ++k) v[k] = 0.0;
  double a[16], b[8], c[4], d[2];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = swap_sum<true>(v[k], v[16 + k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) b[k] = swap_sum<false>(a[k], a[8 + k]);
  {
    const bool hi = lane & 8;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double keep = hi ? b[4 + k] : b[k];
      const double give = hi ? b[k] : b[4 + k];
      c[k] = keep + dpp_full<0x140>(give);  // row_mirror
    }
  }
  {
    const bool hi = lane & 4;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double keep = hi ? c[2 + k] : c[k];
      const double give = hi ? c[k] : c[2 + k];
      d[k] = keep + dpp_full<0x141>(give);  // row_half_mirror
    }
  }
  double e;
  {
    const bool hi = lane & 2;
    const double keep = hi ? d[1] : d[0];
    const double give = hi ? d[0] : d[1];
    e = keep + dpp_full<0x4E>(give);  // quad_perm [2,3,0,1]
  }
  e = e + dpp_full<0xB1>(e);  // quad_perm [1,0,3,2]
  const int q = lane >> 1;
  if ((lane & 1) == 0 && q < KCA3) sm[wave][q] = e;
  __syncthreads();
  if (tid < KCA3) {
    double t = sm[0][tid];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) t += sm[w][tid];
    dst[(size_t)tid * stride] = t;
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
    case PH_OUTER_B: cell_outer_b<T>(C, A, j, i, acc); break;
    case PH_FIN: cell_fin<T>(C, A, j, i, acc); break;
    case PH_CORR_A: cell_corr_a<T>(C, A, j, i, acc); break;
    case PH_CORR_B: cell_corr_b<T>(C, j, i, acc); break;
    case PH_CG0: cell_cg0<T>(C, A, j, i, acc); break;
    case PH_DEFL_A: cell_defl_a(C, j, i, acc); break;
    case PH_DEFL_B: cell_defl_b(C, A, j, i); break;
    case PH_DSTORE: cell_dstore(C, j, i, acc); break;
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

// The phases that run a few times per step, out of line: their code and
// registers stay off the straight-line path of the PCG rounds.
// Cg points at the cavity's descriptor in global memory (a kernel-argument
// copy passed by reference would have to be spilled to scratch first).
template <class T, int V = 0>
__attribute__((noinline)) __device__ void round_cold(int phase, const PfCav* __restrict__ Cg, const PhaseArgs A0, int b,
                                                     double (*sm)[NPS], double* part_out, int pstride) {
  const PfCav& C = *Cg;
  PhaseArgs A = A0;
  // FIN: p = x - x[0] (pressure_solver.py:147-149); the solution is b itself
  // when the right-hand side was zero.
  if (phase == PH_FIN || (phase == PH_CORR_A && C.prec == PF_PREC_MIXED)) A.x0 = A.x_is_b ? C.b[0] : C.x64[0];
  double acc[NP] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < C.rpw; ++k) {
    const int r = b * C.rpw + k;
    if (r >= C.nreg) break;
    const int ty0 = (r / C.tiles_x) * C.tpb;
    const int tyend = min(ty0 + C.tpb, C.tiles_y);
    const int i = (r % C.tiles_x) * TX + threadIdx.x;
    for (int ty = ty0; ty < tyend; ++ty) {
      const int j = ty * TY + threadIdx.y;
      if (i < C.nx && j < C.ny) cell_phase<T>(phase, C, A, j, i, acc);
    }
  }
  if (phase_nparts(phase) > 0) block_partials_phase(phase, acc, sm, part_out, pstride);
}

// block_partials_k<K, 0> with the workgroup totals left in out[k] (LDS)
// instead of the partials array: the same reduction tree, the same bits.
template <int K>
__device__ inline void block_total(double (&v)[NP], double (*sm)[NPS], double* out) {
  const int tid = tid_flat();
  const int lane = tid & 63, wave = tid >> 6;
  wave_reduce_k<K, 0u>(v);
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[wave][k] = v[k];
  }
  __syncthreads();
  if (tid < K) {
    double t = sm[0][tid];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) t = t + sm[w][tid];
    out[tid] = t;
  }
  __syncthreads();
}

// FP64 cavities of one region (PfCav::local; one workgroup holds every cell):
// the CG iterations of a solve run inside the round that starts the solve's
// P phase, instead of one round per P or Q phase. A thread keeps x, r and
// q = A d of its cells in registers (q is formed once, in P; the Q round
// forms it again from the same values, the same bits); the direction d lives
// in LDS, where the stencil reads its neighbours. Each phase is computed as
// the round of that phase computes it: the same cell expressions (dnew_at,
// apply_op5, chunk_q), the same per-thread order of the sums, the same block
// reduction (block_total), the totals folded as the next round folds one
// workgroup's partials (0.0 + t), and the same tails (tail_p, tail_q_sums,
// cg_next's choice). The results are therefore the bits of the
// round-per-phase path. The loop ends after a Q phase whose tail would leave
// the iteration: that Q's partials are stored and the state is left pending
// on Q, as after a Q round, so the next round runs the tail and what follows
// it. Bounded by lin_maxiter. A solve whose residual is NaN runs one round
// per phase instead. MC: cells per thread, at least the region's tiles (4 or
// 8).
template <class T, int MC>
__device__ __attribute__((noinline)) void local_cg(const PfCav* __restrict__ Cg, PfScal& st, double (*sm)[NPS],
                                                   double* tot, double* __restrict__ lD, double* part_out) {
  static_assert(sizeof(T) == sizeof(double), "FP64 cavities only");
  const PfCav& C = *Cg;
  __shared__ double raw[2];
  __shared__ int go;
  // The descriptor's fields, read once (a load from global memory after every
  // tail would sit on the iteration's critical path).
  const int nx = C.nx, ny = C.ny;
  const int nt = C.tiles_y, op = C.op, jac = C.jacobi, maxiter = C.maxiter, nblk = C.nblk;
  const double invd = C.invd;
  const T dx = (T)C.dx, dy = (T)C.dy;
  const int i = threadIdx.x, lj = threadIdx.y;
  const int tid = tid_flat();
  T* X = (T*)C.x;
  T* R = (T*)C.r;
  T* const D0 = (T*)C.d0;
  T* const D1 = (T*)C.d1;
  T xv[MC], rv[MC], qv[MC];
  {
    const T* Dold = (st.it & 1) ? D0 : D1;
#pragma unroll
    for (int t = 0; t < MC; ++t) {
      const int j = t * TY + lj;
      const bool ok = (t < nt) && (i < nx) && (j < ny);
      const int k = ok ? j * nx + i : 0;
      xv[t] = ok ? X[k] : T(0);
      rv[t] = ok ? R[k] : T(0);
      qv[t] = T(0);
      if (ok) lD[k] = Dold[k];  // the previous direction (read back by this thread only)
    }
  }
  auto cell_ok = [&](int t, int& j, int& k) {
    j = t * TY + lj;
    k = j * nx + i;
    return (t < nt) && (i < nx) && (j < ny);
  };
  auto q_at = [&](int j, int k) -> T {
    const T c = lD[k];
    const T e = (i + 1 < nx) ? lD[k + 1] : T(0);
    const T w = (i > 0) ? lD[k - 1] : T(0);
    const T n = (j + 1 < ny) ? lD[k + nx] : T(0);
    const T s = (j > 0) ? lD[k - nx] : T(0);
    return apply_op5<T>(op, c, e, w, n, s, nx, ny, j, i, dx, dy);
  };
  while (true) {
    // P (chunk_p): d = M r (first iteration) or d_old * beta + M r; q = A d.
    const int first = (st.it == 0);
    const double beta = st.beta;
#pragma unroll
    for (int t = 0; t < MC; ++t) {
      int j, k;
      if (cell_ok(t, j, k)) {
        const T z = (T)jacobi_minv(jac, invd, nx, ny, j, i) * rv[t];
        lD[k] = first ? z : (T)((double)lD[k] * beta + (double)z);
      }
    }
    __syncthreads();
    double acc[NP] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < MC; ++t) {
      int j, k;
      if (cell_ok(t, j, k)) {
        qv[t] = q_at(j, k);
        acc[0] += (double)lD[k] * (double)qv[t];
      }
    }
    block_total<1>(acc, sm, raw);  // its barriers also end this phase's LDS reads
    if (tid == 63) {
      tot[0] = 0.0 + raw[0];
      tail_p(st, tot);  // alpha = rho / (d.q); phase Q
    }
    __syncthreads();
    // Q (chunk_q): x += alpha d, r -= alpha q, partials r.(M r) and r.r.
    const double alpha = st.alpha;
    double acq[NP] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < MC; ++t) {
      int j, k;
      if (cell_ok(t, j, k)) {
        xv[t] = (T)((double)xv[t] + alpha * (double)lD[k]);
        rv[t] = (T)((double)rv[t] - alpha * (double)qv[t]);
        const T zw = (T)jacobi_minv(jac, invd, nx, ny, j, i);
        const T z = zw * rv[t];
        acq[0] += (double)rv[t] * (double)z;
        acq[1] += (double)rv[t] * (double)rv[t];
      }
    }
    block_total<2>(acq, sm, raw);
    if (tid == 63) {
      tot[0] = 0.0 + raw[0];
      tot[1] = 0.0 + raw[1];
      // cg_next for FP64 after the Q tail: another P unless the budget is
      // spent or the residual meets the inner tolerance. A NaN residual also
      // leaves the loop: such a solve goes on one round per phase (the same
      // bits), so a host that watches for NaN sees it between rounds.
      const bool more = (st.it_total + 1 < maxiter) && !(sqrt(tot[1]) < st.inner_atol) && !isnan(tot[1]);
      if (more) {
        tail_q_sums(st, tot);
        st.phase = PH_P;  // cg_next's choice when more holds
      }
      go = more ? 1 : 0;
    }
    __syncthreads();
    if (!go) break;
  }
  // Leave the vectors and partials as a Q round leaves them.
  T* Dn = (st.it & 1) ? D1 : D0;  // the last P's direction array
#pragma unroll
  for (int t = 0; t < MC; ++t) {
    int j, k;
    if (cell_ok(t, j, k)) {
      X[k] = xv[t];
      R[k] = rv[t];
      Dn[k] = lD[k];
    }
  }
  if (tid < 2) part_out[(size_t)tid * nblk] = raw[tid];
}

// Body of one round for workgroup b of cavity c. M > 0 (float32 only): the
// cavity's region height in tiles. (Issuing the PCG region loads at entry,
// before the phase is known, gained no time and gave wrong results in
// batches with more workgroups than CUs; they are issued in the PCG branch.)
template <class T, int U, int M, bool XD, bool MULTI>
__device__ __attribute__((always_inline)) inline void round_body(const KArgs& KA, const PfCav& C, const PfCav* Cg, int c, int b, PfScal& st,
                                  double (*sm)[NPS], double* tot, float* lds) {
  const int parity = KA.parity;
  const int tid = tid_flat();
  const uint32_t* src = (const uint32_t*)(KA.scal + (size_t)(parity ^ 1) * KA.max_cav + c);
  uint32_t* dst = (uint32_t*)(KA.scal + (size_t)parity * KA.max_cav + c);
  constexpr int NW32 = (int)(sizeof(PfScal) / 4);
  static_assert(NW32 <= NTH, "state copied with one word per thread");
  // The state words and the previous round's partials (folding waves only:
  // lane l takes workgroups l, l+64, l+128, l+192) are loaded together, before
  // anything is known about the phase. Loads are unconditional (out-of-range
  // lanes read slot 0 and discard it).
  const double* prev = C.part + (size_t)(parity ^ 1) * C.nblk * NPS;
  const int nblk = C.nblk;
  const int lane = tid & 63;
  // Folding waves: wave 0 takes slots 0..2; in the FP64-iterate kernel wave 1
  // takes slots 3..5 (one wave holding all six would need 48 more VGPRs).
  const int fwave = tid >> 6;
  const bool wf = XD ? (fwave < 2) : (fwave == 0);
  const int koff = (XD && fwave == 1) ? KMAX : 0;
  constexpr int PRE = 4;
  double pre[PRE][KMAX];
#pragma unroll
  for (int t = 0; t < PRE; ++t) {
    const int q = lane + 64 * t;
    const size_t qq = (wf && q < nblk) ? (size_t)q : 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) pre[t][k] = prev[(size_t)(koff + k) * nblk + qq];
  }
  const uint32_t sw = src[tid < NW32 ? tid : 0];
  if (tid < NW32) ((uint32_t*)&st)[tid] = sw;
  __syncthreads();
  if (st.pending != PH_IDLE) {
    // The folding waves sum the partials in a fixed order (per lane:
    // increasing workgroup index; then the DPP tree), so every workgroup of
    // the cavity computes the same bits; thread 63 runs the tail. A PCG
    // iteration is pending in nearly every round: its fold (sums only),
    // reduction and tail form one straight-line block.
    const bool hot = (M > 0) && st.pending == PH_CG;
    if (wf) {
      const unsigned mm = hot ? 0u : (phase_maxmask(st.pending) >> koff);
      double pa[NP] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < PRE; ++t)
        if (lane + 64 * t < nblk) {
#pragma unroll
          for (int k = 0; k < KMAX; ++k) pa[k] = redop(mm, k, pa[k], pre[t][k]);
        }
      for (int q = lane + 64 * PRE; q < nblk; q += 64) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k) pa[k] = redop(mm, k, pa[k], prev[(size_t)(koff + k) * nblk + q]);
      }
      if (__builtin_expect(hot, 1))
        wave_reduce_k<KMAX, 0u>(pa);
      else
        wave_reduce_phase(st.pending, pa);
      if (lane == 63) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k) tot[koff + k] = pa[k];
      }
    }
    if constexpr (XD) __syncthreads();  // wave 1's slots visible to thread 63
    if (tid == 63) {
      if (__builtin_expect(hot, 1)) {
        if constexpr (XD)
          tail_cg_xd(C, st, tot);
        else
          tail_cg(C, st, tot);
      } else {
        tail<0>(*Cg, st, tot);
      }
    }
    __syncthreads();  // tail result visible
  }
  const int phase = st.phase;
  if (phase <= PH_IDLE || phase >= PH_DONE) {
    if (b == 0) {
      const int newly_done = st.done_now;
      __syncthreads();
      if (tid == 0) st.done_now = 0;
      __syncthreads();
      if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
      if (tid == 0 && newly_done) atomicSub(KA.active, 1);
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
  A.par = parity;
  A.x0 = (phase == PH_DEFL_B) ? st.dcoef : 0.0;  // FIN reads x[0] in round_cold
  A.cx = C.cxf;
  A.cy = C.cyf;
  A.betaf = (float)A.beta;
  A.alphaf = (float)A.alpha;
  A.minvf = C.invdf;
  A.dset = st.dsetup;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0};
  double* part_out = C.part + (size_t)parity * nblk * NPS + b;
  // Workgroup b owns regions b*rpw .. b*rpw+rpw-1 (fewer at the cavity's end);
  // without MULTI every cavity of the batch has rpw = 1.
  const int r0 = MULTI ? b * C.rpw : b;
  const int r1 = MULTI ? min(r0 + C.rpw, C.nreg) : b + 1;
  if (__builtin_expect((M > 0) && phase == PH_CG, 1)) {
    if constexpr (M > 0) {
      for (int r = r0; r < r1; ++r) {
        if (r > r0) __syncthreads();  // the previous region's LDS reads are done
        RegionIn<M, XD> L;
        region_load<M, XD>(C, parity, r, L);
        region_cg_f32<M, XD>(C, A, r, acc, lds, L);
      }
    }
    block_partials_k<(XD ? KXD : 3), 0u>(acc, sm, part_out, nblk);
  } else if (M == 0 && sizeof(T) == sizeof(double) && phase == PH_P && C.local && !isnan(st.rnorm2)) {
    if constexpr (sizeof(T) == sizeof(double)) {
      if (C.tiles_y <= 4)
        local_cg<T, 4>(Cg, st, sm, tot, (double*)lds, part_out);
      else
        local_cg<T, 8>(Cg, st, sm, tot, (double*)lds, part_out);
    }
    if (b == 0) {
      __syncthreads();
      if (tid == 0) {
        st.pending = PH_Q;
        if (!st.dsetup) st.rounds_run += 1;
      }
      __syncthreads();
      if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
    }
    return;
  } else if (M == 0 && (phase == PH_P || phase == PH_Q)) {
    for (int r = r0; r < r1; ++r) {
      const int ty0 = (r / C.tiles_x) * C.tpb;
      TileIt ti{r % C.tiles_x, ty0, min(ty0 + C.tpb, C.tiles_y)};
      if (phase == PH_P)
        for (int m = 0; m < C.tpb; m += U) chunk_p<T, U>(C, A, ti, acc);
      else
        for (int m = 0; m < C.tpb; m += U) chunk_q<T, U>(C, A, ti, acc);
    }
    if (phase == PH_P)
      block_partials_k<1, 0u>(acc, sm, part_out, nblk);
    else
      block_partials_k<2, 0u>(acc, sm, part_out, nblk);
  } else {
    round_cold<T>(phase, Cg, A, b, sm, part_out, nblk);
  }
  if (b == 0) {
    __syncthreads();
    if (tid == 0) {
      st.pending = phase;
      if (!st.dsetup) st.rounds_run += 1;
    }
    __syncthreads();
    if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
  }
}

// Round body of a two-iteration cavity (PfCav::ca). The pending phase's
// partials are folded by one wave per slot: wave w sums slot w of every
// workgroup (per lane: increasing workgroup index; then the DPP tree), so all
// fourteen slots of the PCG round fold in parallel; thread 63 runs the tail.
template <int M, bool MULTI, bool XD, int V = 0>
__device__ __attribute__((always_inline)) inline void round_body_ca(const KArgs& KA, const PfCav& C, const PfCav* Cg, int c,
                                                                    int b, PfScal& st, double (*sm)[NPS], double* tot,
                                                                    float* lds) {
  const int parity = KA.parity;
  const int tid = tid_flat();
  const uint32_t* src = (const uint32_t*)(KA.scal + (size_t)(parity ^ 1) * KA.max_cav + c);
  uint32_t* dst = (uint32_t*)(KA.scal + (size_t)parity * KA.max_cav + c);
  constexpr int NW32 = (int)(sizeof(PfScal) / 4);
  static_assert(NW32 <= NTH, "state copied with one word per thread");
  static_assert(KCA <= NWAVE, "one folding wave per partial slot");
  const double* prev = C.part + (size_t)(parity ^ 1) * C.nblk * NPS;
  const int nblk = C.nblk;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int PRE = 4;
  double pre[PRE];
#pragma unroll
  for (int t = 0; t < PRE; ++t) {
    const int q = lane + 64 * t;
    pre[t] = prev[(size_t)wave * nblk + (q < nblk ? q : 0)];
  }
  const uint32_t sw = src[tid < NW32 ? tid : 0];
  if (tid < NW32) ((uint32_t*)&st)[tid] = sw;
  __syncthreads();
  if (st.pending != PH_IDLE) {
    const int pend = st.pending;
    const bool hot = pend == PH_CG;
    const int K = hot ? KCA : phase_nparts(pend);
    if (wave < K) {
      const bool mx = !hot && ((phase_maxmask(pend) >> wave) & 1u);
      double v = 0.0;
#pragma unroll
      for (int t = 0; t < PRE; ++t)
        if (lane + 64 * t < nblk) v = mx ? nanmax(v, pre[t]) : v + pre[t];
      for (int q = lane + 64 * PRE; q < nblk; q += 64) {
        const double w = prev[(size_t)wave * nblk + q];
        v = mx ? nanmax(v, w) : v + w;
      }
      v = mx ? wave_reduce<true>(v) : wave_reduce<false>(v);
      if (lane == 63) tot[wave] = v;
    }
    __syncthreads();  // every slot's total visible to thread 63
    if (tid == 63) {
      if (__builtin_expect(hot, 1))
        tail_ca(C, st, tot);
      else
        tail<V>(*Cg, st, tot);
    }
    __syncthreads();  // tail result visible
  }
  const int phase = st.phase;
  if (phase <= PH_IDLE || phase >= PH_DONE) {
    if (b == 0) {
      const int newly_done = st.done_now;
      __syncthreads();
      if (tid == 0) st.done_now = 0;
      __syncthreads();
      if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
      if (tid == 0 && newly_done) atomicSub(KA.active, 1);
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
  A.par = parity;
  A.x0 = (phase == PH_DEFL_B) ? st.dcoef : 0.0;  // FIN reads x[0] in round_cold
  A.cx = C.cxf;
  A.cy = C.cyf;
  A.betaf = (float)A.beta;
  A.alphaf = (float)A.alpha;
  A.minvf = C.invdf;
  A.dset = st.dsetup;
  A.ca0f = (float)st.ca0;
  A.cb0f = (float)st.cb0;
  A.ca1f = (float)st.ca1;
  A.cb1f = (float)st.cb1;
  A.ca_init = (st.ca_inc == 0);
  A.ca0 = st.ca0;
  A.ca1 = st.ca1;
  double acc[NP] = {0.0, 0.0, 0.0, 0.0};
  const int r0 = MULTI ? b * C.rpw : b;
  const int r1 = MULTI ? min(r0 + C.rpw, C.nreg) : b + 1;
  // The partials' address is formed after the region loop (from the
  // descriptor again), which keeps it out of the loop's registers.
  if (__builtin_expect(phase == PH_CG, 1)) {
    for (int r = r0; r < r1; ++r) {
      if (r > r0) __syncthreads();  // the previous region's LDS reads are done
      region_ca<M, XD>(C, A, r, acc, lds);
    }
    block_partials_ca(acc, sm, C.part + (size_t)parity * C.nblk * NPS + b, C.nblk);
  } else {
    round_cold<float, V>(phase, Cg, A, b, sm, C.part + (size_t)parity * C.nblk * NPS + b, C.nblk);
  }
  if (b == 0) {
    __syncthreads();
    if (tid == 0) {
      st.pending = phase;
      if (!st.dsetup) st.rounds_run += 1;
    }
    __syncthreads();
    if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
  }
}

// Round body of a three-iteration cavity (PfCav::ca == 2): as round_body_ca,
// with 27 partial slots (waves fold slots w and w + 16) and tail_ca3.
template <int M, bool MULTI, int V = 0>
__device__ __attribute__((always_inline)) inline void round_body_ca3(const KArgs& KA, const PfCav& C, const PfCav* Cg, int c,
                                                                    int b, PfScal& st, double (*sm)[NPS], double* tot,
                                                                    float* lds) {
  const int parity = KA.parity;
  const int tid = tid_flat();
  const uint32_t* src = (const uint32_t*)(KA.scal + (size_t)(parity ^ 1) * KA.max_cav + c);
  uint32_t* dst = (uint32_t*)(KA.scal + (size_t)parity * KA.max_cav + c);
  constexpr int NW32 = (int)(sizeof(PfScal) / 4);
  static_assert(NW32 <= NTH, "state copied with one word per thread");
  static_assert(KCA3 <= 2 * NWAVE && NPS >= 2 * NWAVE, "two folding slots per wave");
  const double* prev = C.part + (size_t)(parity ^ 1) * C.nblk * NPS;
  const int nblk = C.nblk;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int PRE = 4;
  double pre[PRE], pre2[PRE];  // slots wave and wave + NWAVE
#pragma unroll
  for (int t = 0; t < PRE; ++t) {
    const int q = lane + 64 * t;
    pre[t] = prev[(size_t)wave * nblk + (q < nblk ? q : 0)];
    pre2[t] = prev[(size_t)(wave + NWAVE) * nblk + (q < nblk ? q : 0)];
  }
  const uint32_t sw = src[tid < NW32 ? tid : 0];
  if (tid < NW32) ((uint32_t*)&st)[tid] = sw;
  __syncthreads();
  if (st.pending != PH_IDLE) {
    const int pend = st.pending;
    const bool hot = pend == PH_CG;
    const int K = hot ? KCA3 : phase_nparts(pend);
    if (wave < K) {
      const bool mx = !hot && ((phase_maxmask(pend) >> wave) & 1u);
      double v = 0.0;
#pragma unroll
      for (int t = 0; t < PRE; ++t)
        if (lane + 64 * t < nblk) v = mx ? nanmax(v, pre[t]) : v + pre[t];
      for (int q = lane + 64 * PRE; q < nblk; q += 64) {
        const double w = prev[(size_t)wave * nblk + q];
        v = mx ? nanmax(v, w) : v + w;
      }
      v = mx ? wave_reduce<true>(v) : wave_reduce<false>(v);
      if (lane == 63) tot[wave] = v;
    }
    if (wave + NWAVE < K) {  // the PCG round only: sums
      double v = 0.0;
#pragma unroll
      for (int t = 0; t < PRE; ++t)
        if (lane + 64 * t < nblk) v = v + pre2[t];
      for (int q = lane + 64 * PRE; q < nblk; q += 64) v = v + prev[(size_t)(wave + NWAVE) * nblk + q];
      v = wave_reduce<false>(v);
      if (lane == 63) tot[wave + NWAVE] = v;
    }
    __syncthreads();  // every slot's total visible to thread 63
    if (tid == 63) {
      if (__builtin_expect(hot, 1))
        tail_ca3(C, st, tot);
      else
        tail<V>(*Cg, st, tot);
    }
    __syncthreads();  // tail result visible
  }
  const int phase = st.phase;
  if (phase <= PH_IDLE || phase >= PH_DONE) {
    if (b == 0) {
      const int newly_done = st.done_now;
      __syncthreads();
      if (tid == 0) st.done_now = 0;
      __syncthreads();
      if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
      if (tid == 0 && newly_done) atomicSub(KA.active, 1);
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
  A.par = parity;
  A.x0 = (phase == PH_DEFL_B) ? st.dcoef : 0.0;  // FIN reads x[0] in round_cold
  A.cx = C.cxf;
  A.cy = C.cyf;
  A.betaf = (float)A.beta;
  A.alphaf = (float)A.alpha;
  A.minvf = C.invdf;
  A.dset = st.dsetup;
  A.ca0f = (float)st.ca0;
  A.cb0f = (float)st.cb0;
  A.ca1f = (float)st.ca1;
  A.cb1f = (float)st.cb1;
  A.ca2f = (float)st.ca2;
  A.cb2f = (float)st.cb2;
  A.ca_init = (st.ca_inc == 0);
  A.ca0 = st.ca0;
  A.ca1 = st.ca1;
  double acc[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) acc[k] = 0.0;
  const int r0 = MULTI ? b * C.rpw : b;
  const int r1 = MULTI ? min(r0 + C.rpw, C.nreg) : b + 1;
  // The partials' address is formed after the region loop (from the
  // descriptor again), which keeps it out of the loop's registers.
  if (__builtin_expect(phase == PH_CG, 1)) {
    for (int r = r0; r < r1; ++r) {
      if (r > r0) __syncthreads();  // the previous region's LDS reads are done
      region_ca3<M>(C, A, r, acc, lds);
    }
    block_partials_ca3(acc, sm, C.part + (size_t)parity * C.nblk * NPS + b, C.nblk);
  } else {
    round_cold<float, V>(phase, Cg, A, b, sm, C.part + (size_t)parity * C.nblk * NPS + b, C.nblk);
  }
  if (b == 0) {
    __syncthreads();
    if (tid == 0) {
      st.pending = phase;
      if (!st.dsetup) st.rounds_run += 1;
    }
    __syncthreads();
    if (tid < NW32) dst[tid] = ((const uint32_t*)&st)[tid];
  }
}

// XDK: which FP64-iterate bodies are compiled (bit 0: one PCG iteration per
// round, bit 1: two); 0 when no cavity of the batch keeps its PCG iterate in
// FP64, which keeps the kernel small.
// MULTI: some cavity has several regions per workgroup.
template <class T, int U, bool SINGLE, int XDK, bool MULTI>
__global__ __launch_bounds__(NTH) void k_round(const KArgs KA) {
  __shared__ PfScal st;
  __shared__ double sm[NWAVE][NPS];
  __shared__ double tot[NPS];
  extern __shared__ float lds_dyn[];  // float32 PCG region: region_lds_floats, ca_lds_floats or ca3_lds_floats(tpb) floats
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
  const PfCav* Cg = KA.cavs + c;  // the same descriptor in global memory
  // XCD-aware region order: workgroups are dispatched round-robin over the 8
  // XCDs (g mod 8), each with its own L2. When the cavity's workgroup range is
  // 8-aligned, XCD x gets the contiguous regions [x*nblk/8, (x+1)*nblk/8), so
  // the halo rows/columns a region reads were mostly fetched into the same L2
  // by its neighbours. Partials are indexed by region, so the reduction order
  // does not depend on the mapping.
  const int lg = g - C.blk0;
  const int b = ((C.blk0 & 7) == 0 && (C.nblk & 7) == 0) ? (lg & 7) * (C.nblk >> 3) + (lg >> 3) : lg;
  if constexpr (sizeof(T) == sizeof(float)) {
    if ((XDK & 2) && C.xd && C.ca) {
      switch (C.tpb) {
        case 1: round_body_ca<1, MULTI, true>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 2: round_body_ca<2, MULTI, true>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        default: round_body_ca<4, MULTI, true>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
      }
    } else if ((XDK & 1) && C.xd) {
      switch (C.tpb) {
        case 1: round_body<T, U, 1, true, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 2: round_body<T, U, 2, true, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 4: round_body<T, U, 4, true, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        default: round_body<T, U, 8, true, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
      }
    } else if (C.ca) {
      switch (C.tpb) {
        case 1: round_body_ca<1, MULTI, false>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 2: round_body_ca<2, MULTI, false>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        default: round_body_ca<4, MULTI, false>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
      }
    } else {
      switch (C.tpb) {
        case 1: round_body<T, U, 1, false, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 2: round_body<T, U, 2, false, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        case 4: round_body<T, U, 4, false, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
        default: round_body<T, U, 8, false, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn); break;
      }
    }
  } else {
    round_body<T, U, 0, false, MULTI>(KA, C, Cg, c, b, st, sm, tot, lds_dyn);
  }
}

// Batches whose cavities all run three PCG iterations per round (PfCav::ca
// == 2, regions of 2 or 4 tiles, FP32 iterate). A kernel of its own, so
// that neither its register allocation nor that of the other instances
// pays for the other's code; the cold phases are their own instances (V = 2).
template <bool SINGLE, bool MULTI>
__global__ __launch_bounds__(NTH) void k_round_ca3(const KArgs KA) {
  __shared__ PfScal st;
  __shared__ double sm[NWAVE][NPS];
  __shared__ double tot[NPS];
  extern __shared__ float lds_dyn[];  // ca3_lds_floats(tpb) floats
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
  const PfCav* Cg = KA.cavs + c;
  const int lg = g - C.blk0;
  const int b = ((C.blk0 & 7) == 0 && (C.nblk & 7) == 0) ? (lg & 7) * (C.nblk >> 3) + (lg >> 3) : lg;
  if (C.tpb == 2)
    round_body_ca3<2, MULTI, 2>(KA, C, Cg, c, b, st, sm, tot, lds_dyn);
  else
    round_body_ca3<4, MULTI, 2>(KA, C, Cg, c, b, st, sm, tot, lds_dyn);
}

// Two workgroups per CU: batches whose cavities all run the two-iteration
// FP32-iterate round on regions of at most two tiles, one region per
// workgroup (LDS: 57.6 KB per workgroup). The register budget is 64 VGPRs
// (8 waves per SIMD), so while one workgroup waits on a barrier or a load
// the other computes. The cold phases are their own instances (V = 1),
// compiled within the same budget.
template <bool SINGLE>
__global__ __attribute__((amdgpu_flat_work_group_size(NTH, NTH), amdgpu_waves_per_eu(8, 8))) void k_round_occ2(
    const KArgs KA) {
  __shared__ PfScal st;
  __shared__ double sm[NWAVE][NPS];
  __shared__ double tot[NPS];
  extern __shared__ float lds_dyn[];  // ca_lds_floats(tpb) floats
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
  const PfCav* Cg = KA.cavs + c;
  const int lg = g - C.blk0;
  const int b = ((C.blk0 & 7) == 0 && (C.nblk & 7) == 0) ? (lg & 7) * (C.nblk >> 3) + (lg >> 3) : lg;
  if (C.tpb == 1)
    round_body_ca<1, false, false, 1>(KA, C, Cg, c, b, st, sm, tot, lds_dyn);
  else
    round_body_ca<2, false, false, 1>(KA, C, Cg, c, b, st, sm, tot, lds_dyn);
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

// State set-up of the allocation-free solve: descriptor, both state
// parities, the active counter and the workgroup -> cavity table.
__global__ void k_noalloc_init(const PfCav c, const PfScal s0, PfCav* __restrict__ cav, PfScal* __restrict__ scal,
                               int* __restrict__ active, int* __restrict__ blk, int nblk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    cav[0] = c;
    scal[0] = s0;
    scal[1] = s0;
    active[0] = 1;
  }
  if (t < nblk) blk[t] = 0;
}

// One column of a cavity field as float64 (the driver's centerline).
template <class T>
__global__ void k_column(const T* __restrict__ src, int nx, int ny, int col, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < ny) out[j] = (double)src[(size_t)j * nx + col];
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
