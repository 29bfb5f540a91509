// libpyflow_hip: gfx950 HIP implementation of PyFlow's projection time step.
//
// Execution model
// ---------------
// A batch holds cavities (independent lid-driven-cavity problems of any size).
// Each cavity is cut into 64 x 16-cell tiles, grouped into regions of tpb tiles
// stacked in y, and covered by nblk workgroups (rpw regions each) of
// 64 x 16 threads (one wave per tile row), each owning tpb consecutive tiles.
// Every cavity runs a phase machine on the device:
//
//   STATS -> PRED_A -> PRED_B -> PINIT -> {P -> Q}* [-> OUTER_A -> OUTER_B]*
//         -> FIN -> CORR_A -> CORR_B -> (STATS of the next step | DONE)
//
// A "round" is one launch over all workgroups of all cavities; each executes
// the phase its cavity is in. Phases only read data written in earlier rounds
// (plus their own cells), so no workgroup waits on another inside a launch:
// there is no spin-wait and no memory fence anywhere, and a round always
// drains. Per-cavity reductions are deterministic: each workgroup writes
// fixed-order FP64 partials; in the next round every workgroup of the cavity
// sums them in the same fixed order and runs the cavity's scalar tail (dt
// controller, CG alpha/beta, convergence tests) -- see pf_rounds.hpp.
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
#include <mutex>
#include <string>
#include <vector>

#include "pf_rounds.hpp"
#include "pyflow_hip.h"

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
  int64_t round = 0;  // index of the next round; its parity selects the state buffers
  int unroll = 2;     // tiles per unrolled chunk in the P/Q loops (PYFLOW_UNROLL: 1, 2, 4)
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<PfCav> hcav;       // slot -> description (nx == 0: free)
  std::vector<pf_cavity_params> hpar;  // slot -> the caller's parameters
  std::vector<char> used;
  std::vector<char> dreset;  // slot -> its deflation vector is stale (operator or spacing changed)
  PfCav* dcav = nullptr;
  PfScal* dscal = nullptr;
  int* dtile = nullptr;  // workgroup -> cavity slot
  int* dactive = nullptr;
  PfScal* hscal = nullptr;  // pinned mirror
  int* hactive = nullptr;   // pinned
  int total_blocks = 0;
  int single_slot = -1;  // the only used slot when the batch holds one cavity, else -1
  int max_tpb = 1;       // largest region height (tiles) over the batch's cavities
  int lds_floats = 0;    // dynamic LDS of a float32 round: the largest PCG region's staging
  bool any_xd = false;   // some cavity keeps its PCG iterate in FP64
  int xd_kinds = 0;      // rounds of those cavities: bit 0 one PCG iteration, bit 1 two
  bool multi = false;    // some cavity has several regions per workgroup
  bool occ2 = false;     // every cavity fits the two-workgroups-per-CU kernel
  bool s3 = false;       // every cavity runs three PCG iterations per round (k_round_ca3)
  bool tables_dirty = true;
  int graph_len = 0;                      // rounds per captured graph (even; 0: direct launches)
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // by parity of the first round
  int64_t rounds = 0;
  int64_t predicted_rounds = 0;  // rounds the slowest cavity ran in the last run (first chunk of the next)
  int64_t step_rounds = 0;       // tracked run: rounds per step the cavity ran in the last one
  PfScal* hobs = nullptr;        // pinned: the tracked cavity's state after the pending chunk
  int obs_slot = -1;             // cavity whose step ends size the chunks of this run, else -1
  int64_t max_ahead = 4096;      // tracked run: largest chunk queued on a prediction
  bool track = true;             // size the chunks of one stepping cavity from its step ends (PYFLOW_TRACK)
  int64_t round_guard = 0;
  hipEvent_t ev = nullptr;
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;  // timing of the pending chunk
  bool chunk_pending = false;
  bool scal_fresh = false;  // hscal equals the device's latest state (pulled, nothing launched since)
  bool serving = false;  // chunks queued by pf_batch_advance: poll queues nothing more
  int64_t poll_chunk = 16;  // rounds queued by the next pf_batch_poll (doubles up to max_chunk)
  int64_t max_chunk = 1024;  // largest chunk of rounds between host checks (PYFLOW_MAX_CHUNK)
  bool timing = false;
  double gpu_ms = 0.0;       // summed GPU time of launched chunks (timing on)
  int64_t timed_rounds = 0;
  double* dgather = nullptr;  // device column buffer (pf_batch_centerline), grown on demand
  double* hgather = nullptr;  // pinned host copy of it
  int gather_cap = 0;
  double* hstage = nullptr;  // pinned: two halves of kHalfBytes for field copies (allocated on first use)
  hipEvent_t hev[2] = {nullptr, nullptr};  // copy of each half done
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
  if (p->tile_rows != 0 && p->tile_rows != 1 && p->tile_rows != 2 && p->tile_rows != 4 && p->tile_rows != 8)
    return fail(PF_EINVAL, "tile_rows must be 0, 1, 2, 4 or 8");
  if (p->max_workgroups < 0) return fail(PF_EINVAL, "max_workgroups must be >= 0");
  if (p->pcg_per_round < 0 || p->pcg_per_round > 3) return fail(PF_EINVAL, "pcg_per_round must be 0, 1, 2 or 3");
  if (p->iterate64 < 0 || p->iterate64 > 2) return fail(PF_EINVAL, "iterate64 must be 0, 1 or 2");
  if (p->deflate < 0 || p->deflate > 2) return fail(PF_EINVAL, "deflate must be 0, 1 or 2");
  return PF_OK;
}

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  const int x = atoi(v);
  return x > 0 ? x : dflt;
}

// A switch: unset or empty -> dflt, else the integer value (0 = off).
static int env_switch(const char* name, int dflt) {
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  return atoi(v);
}

// Workgroup regions: M = tpb tiles (64 x 16 cells each) stacked in y, M a
// power of two <= 8, starting from PYFLOW_TPB (default 4) and doubled while a
// cavity would need more than PYFLOW_MAXBLK (default 1024) workgroups.
// Regions of 64 x 16*tpb cells (tpb = 4 unless the cavity is shorter), and
// rpw consecutive regions per workgroup so that a cavity launches at most
// PYFLOW_MAXWG (256: one per CU) workgroups: a workgroup then pays its
// state/partials round trip and tail once per round, not once per region.
// A cavity's tile_rows / max_workgroups override both defaults.
static void blocking_of(const pf_cavity_params* p, int tiles_x, int tiles_y, int* nblk, int* tpb, int* nreg,
                        int* rpw) {
  int m = 1;
  const int target = p->tile_rows > 0 ? p->tile_rows : env_int("PYFLOW_TPB", 4);
  while (m * 2 <= target && m < 8) m *= 2;
  while (m > 1 && m / 2 >= tiles_y) m /= 2;  // no more rows than the cavity has
  const int regions = tiles_x * ((tiles_y + m - 1) / m);
  const int maxwg = std::max(1, p->max_workgroups > 0 ? p->max_workgroups : env_int("PYFLOW_MAXWG", 256));
  int k = 1;
  while ((regions + k - 1) / k > maxwg) k *= 2;
  *tpb = m;
  *nreg = regions;
  *rpw = k;
  *nblk = (regions + k - 1) / k;
}

static size_t cavity_bytes(int prec, const pf_cavity_params* p) {
  const size_t n = (size_t)p->nx * p->ny;
  int tx, ty, nb, tpb;
  tiles_of(p->nx, p->ny, &tx, &ty);
  int nreg, rpw;
  blocking_of(p, tx, ty, &nb, &tpb, &nreg, &rpw);
  size_t s = 0;
  s += 10 * align256(n * esize(prec));  // u v p us vs r d0 d1 x r1
  s += align256(n * sizeof(double));                           // b
  if (prec == PF_PREC_MIXED) s += align256(n * sizeof(double));  // x64
  s += align256((size_t)2 * nb * NPS * sizeof(double));         // partials, two parities
  if (prec == PF_PREC_MIXED) s += align256(n * sizeof(float));   // y (deflation)
  return s;
}

// MIXED: two PCG iterations per round (regions of up to 4 tiles; their
// 4-cell frames fit in LDS), with an FP32 or an FP64 iterate. Default on;
// PYFLOW_CA=0 (FP32 iterate) or PYFLOW_CA_XD=0 (FP64 iterate) turns it off,
// and a cavity's pcg_per_round (1 or 2) overrides both.
// MIXED: three PCG iterations per round (6-cell frames) on FP32-iterate
// cavities with regions of 2 or 4 tiles; PYFLOW_S3=1 makes it the default there.
static void choose_ca(PfCav& c, int prec, int pcg_per_round) {
  const int dflt = c.xd ? (env_switch("PYFLOW_CA_XD", 1) != 0) : (env_switch("PYFLOW_CA", 1) != 0);
  const bool s3_ok = prec == PF_PREC_MIXED && !c.xd && (c.tpb == 2 || c.tpb == 4);
  const int s = pcg_per_round > 0 ? pcg_per_round : (dflt ? (env_switch("PYFLOW_S3", 0) != 0 ? 3 : 2) : 1);
  if (s == 3 && s3_ok)
    c.ca = 2;
  else
    c.ca = (prec == PF_PREC_MIXED && c.tpb <= 4 && s >= 2) ? 1 : 0;
}

// Dynamic LDS of a cavity's round: the float32 PCG region's staging, or in
// FP64 the direction array of the in-round CG (local_cg).
static int lds_floats_of(const PfCav& c) {
  if (c.prec != PF_PREC_MIXED) return c.local ? 2 * c.nx * c.ny : 0;
  return c.ca == 2 ? ca3_lds_floats(c.tpb) : (c.ca ? ca_lds_floats(c.tpb) : region_lds_floats(c.tpb));
}

static void fill_params(PfCav& c, const pf_cavity_params* p, int prec) {
  c.nx = p->nx;
  c.ny = p->ny;
  int tx, ty;
  c.ntiles = tiles_of(p->nx, p->ny, &tx, &ty);
  c.tiles_x = tx;
  c.tiles_y = ty;
  blocking_of(p, tx, ty, &c.nblk, &c.tpb, &c.nreg, &c.rpw);
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
  c.cxf = (float)(1.0 / (p->dx * p->dx));
  c.cyf = (float)(1.0 / (p->dy * p->dy));
  c.invdf = (float)p->inv_diag;
  c.inner_rtol = (p->inner_rtol > 0.0) ? p->inner_rtol : 1e-4;
  // MIXED: accumulate the PCG iterate in FP64 on large grids. Rounding the
  // iterate to FP32 perturbs the true residual by about eps32 * cond(A) of the
  // pass's starting residual; cond(A) ~ (2n/pi)^2 makes that 0.1 at n = 2048
  // and 0.4 at n = 4096, where the refinement passes stop contracting.
  // A cavity's iterate64 (1: FP32, 2: FP64) decides; with 0, PYFLOW_XD=0/1
  // forces it off/on.
  const int xd_env = env_switch("PYFLOW_XD", -1);
  const int nmax = std::max(p->nx, p->ny);
  const int xd_dflt = xd_env >= 0 ? (xd_env ? 1 : 0) : (nmax >= 1536 ? 1 : 0);
  c.xd = (prec == PF_PREC_MIXED) ? (p->iterate64 > 0 ? (p->iterate64 == 2 ? 1 : 0) : xd_dflt) : 0;
  choose_ca(c, prec, p->pcg_per_round);
  // FP64 cavities of one region run a solve's CG iterations inside one round
  // (bitwise the same as one round per phase); PYFLOW_LOCAL=0 turns it off.
  // Its direction array takes nx*ny doubles of LDS, at most 48 KiB here.
  c.local = (prec == PF_PREC_FP64 && c.nreg == 1 && (int64_t)c.nx * c.ny <= 6144 &&
             env_switch("PYFLOW_LOCAL", 1) != 0)
                ? 1
                : 0;
  const int defl_dflt = (p->mode != PF_MODE_SOLVE) ? 1 : 0;
  c.defl = (prec == PF_PREC_MIXED && p->op == PF_OP_NEG_LAP_PINNED)
               ? (p->deflate > 0 ? (p->deflate == 2 ? 1 : 0) : defl_dflt)
               : 0;
}

// y = A^-1 1 depends on the operator and the spacing: a cavity whose
// parameters change them solves it again before its next solve.
static bool same_operator(const PfCav& a, const PfCav& b) { return a.op == b.op && a.dx == b.dx && a.dy == b.dy; }

static void drop_graphs(pf_batch* b);

static int sync_tables(pf_batch* b) {
  if (!b->tables_dirty) return PF_OK;
  drop_graphs(b);
  std::vector<int> tiles;  // workgroup -> cavity
  tiles.reserve(b->max_tiles);
  for (int s = 0; s < b->max_cav; ++s) {
    if (!b->used[s]) continue;
    PfCav& c = b->hcav[s];
    c.blk0 = (int)tiles.size();
    for (int t = 0; t < c.nblk; ++t) tiles.push_back(s);
  }
  if ((int64_t)tiles.size() > b->max_tiles) return fail(PF_EINVAL, "batch capacity exceeded");
  b->total_blocks = (int)tiles.size();
  int nused = 0, last = -1;
  for (int s = 0; s < b->max_cav; ++s)
    if (b->used[s]) {
      ++nused;
      last = s;
    }
  b->single_slot = (nused == 1) ? last : -1;
  b->max_tpb = 1;
  b->lds_floats = 0;
  b->any_xd = false;
  b->xd_kinds = 0;
  b->multi = false;
  b->occ2 = (b->prec == PF_PREC_MIXED) && env_switch("PYFLOW_OCC2", 1) != 0;
  int n3 = 0;
  for (int s = 0; s < b->max_cav; ++s)
    if (b->used[s] && b->hcav[s].ca == 2) ++n3;
  if (n3 > 0 && n3 < nused)
    return fail(PF_EINVAL, "cavities with three PCG iterations per round (pcg_per_round 3) cannot share a batch "
                           "with cavities that run one or two");
  b->s3 = n3 > 0;
  for (int s = 0; s < b->max_cav; ++s)
    if (b->used[s]) {
      b->max_tpb = std::max(b->max_tpb, b->hcav[s].tpb);
      b->lds_floats = std::max(b->lds_floats, lds_floats_of(b->hcav[s]));
      b->any_xd = b->any_xd || b->hcav[s].xd;
      if (b->hcav[s].xd) b->xd_kinds |= b->hcav[s].ca ? 2 : 1;
      const PfCav& c = b->hcav[s];
      b->occ2 = b->occ2 && c.ca == 1 && !c.xd && c.tpb <= 2 && c.rpw == 1;
      b->multi = b->multi || b->hcav[s].rpw > 1;
    }
  HIPCHK(hipMemcpyAsync(b->dcav, b->hcav.data(), sizeof(PfCav) * b->max_cav, hipMemcpyHostToDevice, b->stream));
  if (!tiles.empty())
    HIPCHK(hipMemcpyAsync(b->dtile, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tables_dirty = false;
  return PF_OK;
}

// Host edits of a cavity's state go to both parity buffers (the batch is idle).
static int push_scal(pf_batch* b, int slot) {
  for (int par = 0; par < 2; ++par)
    HIPCHK(hipMemcpyAsync(b->dscal + (size_t)par * b->max_cav + slot, b->hscal + slot, sizeof(PfScal),
                          hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

static int push_scal_all(pf_batch* b) {
  for (int par = 0; par < 2; ++par)
    HIPCHK(hipMemcpyAsync(b->dscal + (size_t)par * b->max_cav, b->hscal, sizeof(PfScal) * b->max_cav,
                          hipMemcpyHostToDevice, b->stream));
  return PF_OK;
}

// The latest state is in the buffer written by the last launched round.
static int pull_scal(pf_batch* b) {
  const int par = (int)((b->round + 1) & 1);
  HIPCHK(hipMemcpyAsync(b->hscal, b->dscal + (size_t)par * b->max_cav, sizeof(PfScal) * b->max_cav,
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->scal_fresh = true;
  return PF_OK;
}

static int check_slot(pf_batch* b, int slot) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  if (slot < 0 || slot >= b->max_cav || !b->used[slot]) return fail(PF_EINVAL, "bad cavity slot");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  return PF_OK;
}

// The kernel instance a round of this batch runs, with its launch shape.
struct RoundLaunch {
  const void* fn = nullptr;
  dim3 grid, blk;
  size_t lds = 0;
};

template <class T, int U, bool MULTI>
static const void* round_fn_m(const pf_batch* b) {
  // FP64-iterate cavities are large (a round takes 100+ us): their descriptor
  // is read from the table (the by-value copy would only add register
  // pressure to the bigger kernel).
  if (sizeof(T) == sizeof(float) && b->any_xd) {
    if (b->xd_kinds == 2) return (const void*)k_round<T, U, false, 2, MULTI>;
    if (b->xd_kinds == 3) return (const void*)k_round<T, U, false, 3, MULTI>;
    return (const void*)k_round<T, U, false, 1, MULTI>;
  }
  if (b->single_slot >= 0) return (const void*)k_round<T, U, true, 0, MULTI>;
  return (const void*)k_round<T, U, false, 0, MULTI>;
}

template <class T, int U>
static const void* round_fn_u(const pf_batch* b) {
  return b->multi ? round_fn_m<T, U, true>(b) : round_fn_m<T, U, false>(b);
}

static RoundLaunch round_launch(const pf_batch* b) {
  RoundLaunch L;
  L.blk = dim3(TX, TY);
  L.grid = dim3(b->total_blocks);
  if (b->prec == PF_PREC_MIXED) {
    if (b->s3) {
      const bool one = b->single_slot >= 0;
      L.fn = b->multi ? (one ? (const void*)k_round_ca3<true, true> : (const void*)k_round_ca3<false, true>)
                      : (one ? (const void*)k_round_ca3<true, false> : (const void*)k_round_ca3<false, false>);
    } else if (b->occ2)
      L.fn = b->single_slot >= 0 ? (const void*)k_round_occ2<true> : (const void*)k_round_occ2<false>;
    else
      L.fn = round_fn_u<float, 2>(b);  // U only shapes the FP64 P/Q loops
    L.lds = (size_t)b->lds_floats * sizeof(float);
  } else {
    L.fn = b->unroll >= 4 ? round_fn_u<double, 4>(b) : b->unroll == 2 ? round_fn_u<double, 2>(b)
                                                                      : round_fn_u<double, 1>(b);
    L.lds = (size_t)b->lds_floats * sizeof(float);
  }
  return L;
}

static KArgs round_args(const pf_batch* b, int parity) {
  KArgs K;
  std::memset(&K, 0, sizeof(K));
  K.cavs = b->dcav;
  K.scal = b->dscal;
  K.blk_cav = b->dtile;
  K.active = b->dactive;
  K.max_cav = b->max_cav;
  K.parity = parity;
  K.slot0 = b->single_slot;
  if (b->single_slot >= 0) K.cav0 = b->hcav[b->single_slot];
  return K;
}

static hipError_t launch_rounds_direct(pf_batch* b, int64_t n) {
  if (b->total_blocks <= 0 || n <= 0) return hipSuccess;
  const RoundLaunch L = round_launch(b);
  for (int64_t r = 0; r < n; ++r) {
    KArgs K = round_args(b, (int)(b->round & 1));
    void* args[] = {&K};
    const hipError_t e = hipLaunchKernel(L.fn, L.grid, L.blk, args, L.lds, b->stream);
    if (e != hipSuccess) return e;
    b->round += 1;
  }
  return hipSuccess;
}

static void drop_graphs(pf_batch* b) {
  for (int par = 0; par < 2; ++par)
    if (b->gexec[par]) {
      (void)hipGraphExecDestroy(b->gexec[par]);
      b->gexec[par] = nullptr;
    }
}

// A chain of graph_len rounds whose first round has parity `par`: kernel
// nodes only (each round depends on the one before), so a replay queues
// graph_len rounds with one host call. The kernel arguments are fixed when
// the graph is built; any change of the batch's tables drops the graphs.
static int build_graph(pf_batch* b, int par) {
  const RoundLaunch L = round_launch(b);
  hipGraph_t g = nullptr;
  HIPCHK(hipGraphCreate(&g, 0));
  hipGraphNode_t prev = nullptr;
  hipError_t e = hipSuccess;
  for (int k = 0; k < b->graph_len && e == hipSuccess; ++k) {
    KArgs K = round_args(b, par ^ (k & 1));
    void* args[] = {&K};
    hipKernelNodeParams np;
    std::memset(&np, 0, sizeof(np));
    np.func = const_cast<void*>(L.fn);
    np.gridDim = L.grid;
    np.blockDim = L.blk;
    np.sharedMemBytes = (unsigned)L.lds;
    np.kernelParams = args;
    np.extra = nullptr;
    hipGraphNode_t node = nullptr;
    e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &np);
    prev = node;
  }
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(PF_EHIP, std::string("graph build failed: ") + hipGetErrorString(e));
  b->gexec[par] = ex;
  return PF_OK;
}

// Queues n rounds: whole graph replays where graphs are on, the rest (and
// everything when they are off) as direct launches. Either way the same
// kernels run with the same arguments in the same order.
static int launch_rounds(pf_batch* b, int64_t n) {
  if (b->total_blocks > 0 && b->graph_len > 0 && n >= b->graph_len) {
    const int par = (int)(b->round & 1);
    if (!b->gexec[par]) {
      const int rc = build_graph(b, par);
      if (rc) return rc;
    }
    const int64_t reps = n / b->graph_len;
    for (int64_t k = 0; k < reps; ++k) HIPCHK(hipGraphLaunch(b->gexec[par], b->stream));
    b->round += reps * b->graph_len;  // graph_len is even: the parity is unchanged
    b->rounds += reps * b->graph_len;
    n -= reps * b->graph_len;
  }
  HIPCHK(launch_rounds_direct(b, n));
  b->rounds += n;
  return PF_OK;
}

static void diag_from(const PfScal& s, pf_step_diag* d) {
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
  d->rp_iterations_run = s.it_run;
  d->outer_passes_run = s.outer_run;
}

static void fill_diag(const pf_batch* b, int slot, pf_step_diag* d) { diag_from(b->hscal[slot], d); }

// Host <-> device field copies between runs. FP64 fields are copied in
// place; MIXED fields convert through an FP64 staging area, the cavity's
// predictor pair u*, v* (contiguous, 2 x align256(4 n) >= 8 n bytes). The
// predictor pair is written by PRED_A and dead once a step's CORR_B has run,
// so between runs it holds nothing: staging through it leaves the Poisson
// right-hand side b (pf_batch_set_rhs) and the solution untouched.
static double* stage_of(PfCav& c) { return (double*)c.us; }

// Fields cross between host and device through two pinned halves: the host
// packs (or unpacks) one chunk of rows, dropping the ghost ring, while the
// previous chunk's copy runs. A pageable 2-D copy (hipMemcpy2DAsync with a
// host row pitch) moved about 0.7 GB/s here: 35 ms for the three 1024^2
// fields of a drop-in upload. Rows longer than a half take the 2-D copy.
constexpr size_t kHalfBytes = size_t(8) << 20;

static int ensure_hstage(pf_batch* b) {
  if (b->hstage) return PF_OK;
  HIPCHK(hipHostMalloc((void**)&b->hstage, 2 * kHalfBytes, hipHostMallocDefault));
  for (int k = 0; k < 2; ++k)
    if (!b->hev[k]) HIPCHK(hipEventCreateWithFlags(&b->hev[k], hipEventDisableTiming));
  return PF_OK;
}

static bool rows_fit(int nx) { return (size_t)nx * sizeof(double) <= kHalfBytes; }

// ny host rows of nx doubles, ld apart -> the compact device array dst.
// Queued on the batch stream; the caller synchronizes.
static int rows_h2d(pf_batch* b, double* dst, const double* src, int64_t ld, int nx, int ny) {
  int rc = ensure_hstage(b);
  if (rc) return rc;
  const int64_t cr = (int64_t)(kHalfBytes / (sizeof(double) * nx));
  bool used[2] = {false, false};
  int k = 0;
  for (int64_t j0 = 0; j0 < ny; j0 += cr, k ^= 1) {
    const int64_t n = std::min<int64_t>(cr, ny - j0);
    double* h = b->hstage + k * (kHalfBytes / sizeof(double));
    if (used[k]) HIPCHK(hipEventSynchronize(b->hev[k]));  // this half's last copy has run
    if (ld == nx) {
      std::memcpy(h, src + j0 * ld, sizeof(double) * n * nx);
    } else {
      for (int64_t j = 0; j < n; ++j) std::memcpy(h + j * nx, src + (j0 + j) * ld, sizeof(double) * nx);
    }
    HIPCHK(hipMemcpyAsync(dst + j0 * nx, h, sizeof(double) * n * nx, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(b->hev[k], b->stream));
    used[k] = true;
  }
  return PF_OK;
}

// The compact device array src (ny rows of nx doubles) -> host rows ld apart.
// Returns with every row written.
static int rows_d2h(pf_batch* b, double* dst, const double* src, int64_t ld, int nx, int ny) {
  int rc = ensure_hstage(b);
  if (rc) return rc;
  const int64_t cr = (int64_t)(kHalfBytes / (sizeof(double) * nx));
  auto unpack = [&](int k, int64_t j0) {
    const int64_t n = std::min<int64_t>(cr, ny - j0);
    const double* h = b->hstage + k * (kHalfBytes / sizeof(double));
    if (ld == nx) {
      std::memcpy(dst + j0 * ld, h, sizeof(double) * n * nx);
    } else {
      for (int64_t j = 0; j < n; ++j) std::memcpy(dst + (j0 + j) * ld, h + j * nx, sizeof(double) * nx);
    }
  };
  int k = 0;
  int64_t prev = -1;  // first row of the chunk whose copy went to half k ^ 1
  for (int64_t j0 = 0; j0 < ny; j0 += cr, k ^= 1) {
    const int64_t n = std::min<int64_t>(cr, ny - j0);
    double* h = b->hstage + k * (kHalfBytes / sizeof(double));
    HIPCHK(hipMemcpyAsync(h, src + j0 * nx, sizeof(double) * n * nx, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipEventRecord(b->hev[k], b->stream));
    if (prev >= 0) {
      HIPCHK(hipEventSynchronize(b->hev[k ^ 1]));
      unpack(k ^ 1, prev);
    }
    prev = j0;
  }
  if (prev >= 0) {
    HIPCHK(hipEventSynchronize(b->hev[k ^ 1]));
    unpack(k ^ 1, prev);
  }
  return PF_OK;
}

template <class T>
static int upload_field(pf_batch* b, PfCav& c, T* dst, const double* src, int64_t ld, int ghost) {
  const size_t n = (size_t)c.nx * c.ny;
  const double* s0 = src + (ghost ? (ld + 1) : 0);
  double* to = (sizeof(T) == sizeof(double)) ? (double*)dst : stage_of(c);
  if (rows_fit(c.nx)) {
    int rc = rows_h2d(b, to, s0, ld, c.nx, c.ny);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpy2DAsync(to, sizeof(double) * c.nx, s0, sizeof(double) * ld, sizeof(double) * c.nx, c.ny,
                            hipMemcpyHostToDevice, b->stream));
  }
  if (sizeof(T) != sizeof(double))
    k_f64_to<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream>>>(to, dst, n);
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

template <class T>
static int download_field(pf_batch* b, PfCav& c, const T* src, double* dst, int64_t ld, int ghost) {
  const size_t n = (size_t)c.nx * c.ny;
  double* d0 = dst + (ghost ? (ld + 1) : 0);
  const double* from = (const double*)src;
  if (sizeof(T) != sizeof(double)) {
    double* stage = stage_of(c);
    k_to_f64<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream>>>(src, stage, n);
    from = stage;
  }
  if (rows_fit(c.nx)) {
    int rc = rows_d2h(b, d0, from, ld, c.nx, c.ny);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpy2DAsync(d0, sizeof(double) * ld, from, sizeof(double) * c.nx, sizeof(double) * c.nx, c.ny,
                            hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free solve (pf_solve_noalloc): every buffer of the solve except
// the caller's right-hand side and solution lives in this array. Each GPU the
// library runs on holds its own copy for the life of the process; its device
// address is looked up once per device (hipGetSymbolAddress) and kept.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[PF_NOALLOC_SCRATCH_BYTES];

constexpr int PF_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[PF_MAX_DEVICES] = {};

// The current device's address of `scratch` (resolved on its first use).
static int scratch_for_current_device(char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= PF_MAX_DEVICES) return fail(PF_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(PF_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = (char*)g_scratch_ptr[dev];
  return PF_OK;
}

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

const char* pf_last_error(void) { return g_err.c_str(); }

size_t pf_sizeof_params(void) { return sizeof(pf_cavity_params); }
size_t pf_sizeof_diag(void) { return sizeof(pf_step_diag); }

int pf_device_init(int device) {
  HIPCHK(hipSetDevice(device));
  char* scr = nullptr;  // resolve this device's allocation-free scratch address now
  return scratch_for_current_device(&scr);
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

// The float32 PCG regions stage up to 5 x 72 x 72 floats (101 KiB; the
// two-iteration round on 4 tiles) in LDS, above the 64 KiB a launch may
// request without opting in.
// Rounds per graph: even (a replay then keeps the parity sequence), 0 = off.
static int graph_len_of(int rounds) { return rounds <= 0 ? 0 : std::min(1024, std::max(2, rounds & ~1)); }

static int allow_region_lds() {
  const int bytes = std::max({region_lds_floats(8), ca_lds_floats(4), ca3_lds_floats(4)}) * (int)sizeof(float);
  const void* fns[16] = {
      (const void*)k_round_ca3<true, false>, (const void*)k_round_ca3<false, false>,
      (const void*)k_round_ca3<true, true>, (const void*)k_round_ca3<false, true>,
      (const void*)k_round_occ2<true>, (const void*)k_round_occ2<false>,
      (const void*)k_round<float, 2, true, 0, false>,  (const void*)k_round<float, 2, false, 0, false>,
      (const void*)k_round<float, 2, false, 1, false>, (const void*)k_round<float, 2, false, 2, false>,
      (const void*)k_round<float, 2, false, 3, false>, (const void*)k_round<float, 2, true, 0, true>,
      (const void*)k_round<float, 2, false, 0, true>,  (const void*)k_round<float, 2, false, 1, true>,
      (const void*)k_round<float, 2, false, 2, true>,  (const void*)k_round<float, 2, false, 3, true>};
  for (const void* f : fns) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  return PF_OK;
}

int pf_batch_create_empty(int precision, int max_cav, int64_t max_tiles, void* stream, pf_batch** out) {
  if (!out || max_cav <= 0 || max_tiles <= 0) return fail(PF_EINVAL, "bad arguments");
  if (precision != PF_PREC_FP64 && precision != PF_PREC_MIXED) return fail(PF_EINVAL, "bad precision");
  if (precision == PF_PREC_MIXED) {
    const int rc = allow_region_lds();
    if (rc) return rc;
  }
  pf_batch* b = new pf_batch();
  b->prec = precision;
  b->max_cav = max_cav;
  b->max_tiles = max_tiles;
  b->stream = (hipStream_t)stream;
  b->hcav.assign(max_cav, PfCav{});
  b->hpar.assign(max_cav, pf_cavity_params{});
  b->unroll = env_int("PYFLOW_UNROLL", 2);
  b->graph_len = graph_len_of(env_switch("PYFLOW_GRAPH", 64));
  // Untracked runs (batches, polls) double their chunks up to max_chunk rounds
  // between host checks: each check idles the GPU while the host queues the
  // next chunk, and the last chunk of a run overshoots by at most its length
  // (empty rounds). 1024 against 256 made the 200 steps of the 1024^2 cavity
  // 0.6% faster before that run was tracked. A tracked run (one cavity, several
  // steps) queues at most max_ahead rounds on a prediction (tracked_chunk).
  b->max_chunk = std::max<int64_t>(16, env_int("PYFLOW_MAX_CHUNK", 1024));
  b->track = env_switch("PYFLOW_TRACK", 1) != 0;
  b->max_ahead = std::max<int64_t>(16, env_int("PYFLOW_MAX_AHEAD", 4096));
  b->used.assign(max_cav, 0);
  b->dreset.assign(max_cav, 0);
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dcav, sizeof(PfCav) * max_cav);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dscal, 2 * sizeof(PfScal) * max_cav);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dtile, sizeof(int) * max_tiles);
  if (e == hipSuccess) e = hipMalloc((void**)&b->dactive, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hscal, sizeof(PfScal) * max_cav, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hactive, sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->hobs, sizeof(PfScal), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_t0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_t1);
  if (e != hipSuccess) {
    std::string msg = std::string("allocation failed: ") + hipGetErrorString(e);
    pf_batch_destroy(b);
    return fail(PF_ENOMEM, msg);
  }
  std::memset(b->hscal, 0, sizeof(PfScal) * max_cav);
  *b->hactive = 0;
  e = hipMemsetAsync(b->dscal, 0, 2 * sizeof(PfScal) * max_cav, b->stream);
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
  int64_t blocks_now = 0;
  for (int k = 0; k < b->max_cav; ++k)
    if (b->used[k]) blocks_now += b->hcav[k].nblk;
  if (blocks_now + c.nblk > b->max_tiles) return fail(PF_ENOMEM, "batch capacity exceeded");
  const size_t n = (size_t)c.nx * c.ny;
  const size_t vb = align256(n * esize(b->prec));
  char* w = (char*)workspace;
  void** vecs[10] = {&c.u, &c.v, &c.p, &c.us, &c.vs, &c.r, &c.d0, &c.d1, &c.x, &c.r1};
  for (int k = 0; k < 10; ++k) {
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
  w += align256((size_t)2 * c.nblk * NPS * sizeof(double));
  c.y = (b->prec == PF_PREC_MIXED) ? (float*)w : nullptr;
  HIPCHK(hipMemsetAsync(workspace, 0, need, b->stream));
  b->hcav[s] = c;
  b->hpar[s] = *params;
  b->used[s] = 1;
  b->dreset[s] = 0;
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
  b->hpar[slot] = pf_cavity_params{};
  b->dreset[slot] = 0;
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
  drop_graphs(b);
  if (b->dcav) (void)hipFree(b->dcav);
  if (b->dscal) (void)hipFree(b->dscal);
  if (b->dtile) (void)hipFree(b->dtile);
  if (b->dactive) (void)hipFree(b->dactive);
  if (b->hscal) (void)hipHostFree(b->hscal);
  if (b->hactive) (void)hipHostFree(b->hactive);
  if (b->hobs) (void)hipHostFree(b->hobs);
  if (b->ev) (void)hipEventDestroy(b->ev);
  if (b->ev_t0) (void)hipEventDestroy(b->ev_t0);
  if (b->ev_t1) (void)hipEventDestroy(b->ev_t1);
  if (b->dgather) (void)hipFree(b->dgather);
  if (b->hgather) (void)hipHostFree(b->hgather);
  if (b->hstage) (void)hipHostFree(b->hstage);
  if (b->hev[0]) (void)hipEventDestroy(b->hev[0]);
  if (b->hev[1]) (void)hipEventDestroy(b->hev[1]);
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
  // The workspace (partials: 2 x nblk slots) was carved for the blocking
  // chosen at creation; keep it whatever the environment says now.
  PfCav nc;
  std::memcpy(&nc, &c, sizeof(PfCav));
  fill_params(nc, params, b->prec);
  nc.nblk = c.nblk;
  nc.tpb = c.tpb;
  nc.nreg = c.nreg;
  nc.rpw = c.rpw;
  nc.local = c.local;  // sized into the batch's LDS with that blocking
  choose_ca(nc, b->prec, params->pcg_per_round);
  b->hpar[slot] = *params;
  if (!same_operator(nc, c)) b->dreset[slot] = 1;
  // Unchanged parameters (the drop-in step loop sets them before every step)
  // leave the device tables and the instantiated graphs as they are.
  if (std::memcmp(&nc, &c, sizeof(PfCav)) != 0) {
    std::memcpy(&c, &nc, sizeof(PfCav));
    b->tables_dirty = true;
  }
  return PF_OK;
}

int pf_batch_resolved_params(pf_batch* b, int slot, pf_cavity_params* out) {
  if (!out) return fail(PF_EINVAL, "out is NULL");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  const PfCav& c = b->hcav[slot];
  *out = b->hpar[slot];
  out->tile_rows = c.tpb;
  out->max_workgroups = c.nblk;
  out->pcg_per_round = (b->prec == PF_PREC_MIXED) ? c.ca + 1 : 1;
  out->iterate64 = (b->prec == PF_PREC_MIXED) ? (c.xd ? 2 : 1) : 0;
  out->deflate = (b->prec == PF_PREC_MIXED) ? (c.defl ? 2 : 1) : 0;
  out->inner_rtol = c.inner_rtol;
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
  // keep the last 5 entries, as hist[-5:] does
  const int n = hist_len > 5 ? 5 : hist_len;
  double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < n; ++k) h[k] = hist[hist_len - n + k];
  // The drop-in step loop passes back what the last step left on the device:
  // then there is nothing to upload.
  if (b->scal_fresh && s.dt_prev == dt_prev && s.hist_len == n && std::memcmp(s.hist, h, sizeof(h)) == 0)
    return PF_OK;
  s.dt_prev = dt_prev;
  std::memcpy(s.hist, h, sizeof(h));
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
  if (rows_fit(c.nx)) {
    rc = rows_h2d(b, c.b, rhs, ld, c.nx, c.ny);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpy2DAsync(c.b, sizeof(double) * c.nx, rhs, sizeof(double) * ld, sizeof(double) * c.nx, c.ny,
                            hipMemcpyHostToDevice, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int pf_batch_get_solution(pf_batch* b, int slot, double* x, int64_t ld) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  PfCav& c = b->hcav[slot];
  if (!x || ld < c.nx) return fail(PF_EINVAL, "bad output");
  const double* src = b->hscal[slot].x_is_b ? c.b : c.x64;
  if (rows_fit(c.nx)) {
    rc = rows_d2h(b, x, src, ld, c.nx, c.ny);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpy2DAsync(x, sizeof(double) * ld, src, sizeof(double) * c.nx, sizeof(double) * c.nx, c.ny,
                            hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

static int arm(pf_batch* b, int nsteps, const int32_t* active) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  if (nsteps <= 0) return fail(PF_EINVAL, "nsteps must be > 0");
  int rc = sync_tables(b);
  if (rc) return rc;
  b->obs_slot = -1;
  int nact = 0;
  int64_t guard = 64;
  for (int s = 0; s < b->max_cav; ++s) {
    PfScal& sc = b->hscal[s];
    if (!b->used[s]) continue;
    if (b->dreset[s]) {
      sc.dready = 0;
      b->dreset[s] = 0;
    }
    const bool on = !active || active[s];
    sc.pending = PH_IDLE;
    sc.done_now = 0;
    if (!on) {
      sc.phase = PH_IDLE;
      continue;
    }
    const PfCav& c = b->hcav[s];
    sc.steps_done = 0;
    sc.it_run = 0;
    sc.outer_run = 0;
    sc.rounds_run = 0;
    sc.step_round = 0;
    sc.step_len = 0;
    sc.dsetup = 0;
    if (c.mode == PF_MODE_STEP) {
      sc.phase = PH_STATS;
      sc.steps_left = nsteps;
    } else if (c.mode == PF_MODE_PROJECT) {
      sc.phase = PH_PRED_A;
      sc.steps_left = 1;
    } else {
      enter_pinit(c, sc);
      sc.steps_left = 1;
    }
    const int64_t per = 4 * (int64_t)std::max(c.maxiter, 0) + 80;
    const int64_t setup = c.defl ? 2 * (int64_t)std::max(c.maxiter, 0) + 80 : 0;  // deflation set-up solve
    guard = std::max(guard, (int64_t)sc.steps_left * per + setup + 64);
    ++nact;
  }
  *b->hactive = nact;
  int rc2 = push_scal_all(b);
  if (rc2) return rc2;
  HIPCHK(hipMemcpyAsync(b->dactive, b->hactive, sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->rounds = 0;
  b->round_guard = guard;
  return PF_OK;
}

static int chunk(pf_batch* b, int64_t n) {
  b->scal_fresh = false;
  if (b->timing) HIPCHK(hipEventRecord(b->ev_t0, b->stream));
  const int rc = launch_rounds(b, n);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  if (b->timing) {
    HIPCHK(hipEventRecord(b->ev_t1, b->stream));
    b->timed_rounds += n;
  }
  HIPCHK(hipMemcpyAsync(b->hactive, b->dactive, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  if (b->obs_slot >= 0) {
    const int par = (int)((b->round + 1) & 1);  // buffer of the last round
    HIPCHK(hipMemcpyAsync(b->hobs, b->dscal + (size_t)par * b->max_cav + b->obs_slot, sizeof(PfScal),
                          hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipEventRecord(b->ev, b->stream));
  b->chunk_pending = true;
  return PF_OK;
}

// Rounds to queue next in a tracked run (one cavity stepping nsteps > 1 steps),
// from its state after the last chunk: the length of its last step gives the
// rounds left, and the chunk queues half of them, so the run needs few host
// checks and ends with few empty rounds even where step lengths drift (from
// rest they change several-fold within a few steps). Before the first step
// end, and once the prediction is used up, the chunks double from 16.
static int64_t tracked_chunk(const pf_batch* b, int64_t& over, int nsteps) {
  const PfScal& o = *b->hobs;
  if (o.steps_done > 0) {
    const int64_t left = (int64_t)(nsteps - o.steps_done) * o.step_len - (o.rounds_run - o.step_round);
    if (left > 32) {
      over = 0;
      return std::min<int64_t>(b->max_ahead, left / 2);
    }
  }
  over = over ? std::min<int64_t>(2 * over, b->max_chunk) : 16;
  return over;
}

int pf_batch_run(pf_batch* b, int nsteps, const int32_t* active, pf_step_diag* diag) {
  if (b) b->serving = false;
  int rc = arm(b, nsteps, active);
  if (rc) return rc;
  const int os = b->single_slot;
  const bool stepping = os >= 0 && b->hcav[os].mode == PF_MODE_STEP;
  if (*b->hactive > 0) {
    b->obs_slot = (b->track && stepping && nsteps > 1 && *b->hactive == 1) ? os : -1;
    // First chunk: 15/16 of the rounds the last run predicts for this one
    // (half of them in a tracked run, whose later chunks follow the steps).
    int64_t first = b->predicted_rounds;
    if (stepping && b->step_rounds > 0) first = std::min<int64_t>(first, b->step_rounds * nsteps);
    first = (first * 15) / 16;
    if (b->obs_slot >= 0) first = std::min<int64_t>(b->max_ahead, b->step_rounds * nsteps / 2);
    first = std::max<int64_t>(8, first);
    int64_t step = 16;
    int64_t over = 0;  // tracked run: last doubling chunk
    rc = chunk(b, first);
    if (rc) return rc;
    while (true) {
      HIPCHK(hipEventSynchronize(b->ev));
      b->chunk_pending = false;
      if (b->timing) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev_t0, b->ev_t1));
        b->gpu_ms += ms;
      }
      if (*b->hactive <= 0) break;
      if (b->rounds > b->round_guard) return fail(PF_ESTATE, "round limit exceeded (internal error)");
      const int64_t n = (b->obs_slot >= 0) ? tracked_chunk(b, over, nsteps) : step;
      rc = chunk(b, n);
      if (rc) return rc;
      step = std::min<int64_t>(step * 2, b->max_chunk);
    }
  }
  b->obs_slot = -1;
  rc = pull_scal(b);
  if (rc) return rc;
  if (*b->hactive <= 0) {
    // The next run's first chunk: the rounds the slowest cavity needed this
    // time (the launched rounds include the last chunk's overshoot), and for
    // one stepping cavity its rounds per step.
    int64_t need = 0;
    for (int s = 0; s < b->max_cav; ++s)
      if (b->used[s]) need = std::max<int64_t>(need, b->hscal[s].rounds_run);
    if (need > 0) b->predicted_rounds = need;
    if (stepping && b->hscal[os].steps_done > 0) b->step_rounds = b->hscal[os].rounds_run / b->hscal[os].steps_done;
  }
  if (diag)
    for (int s = 0; s < b->max_cav; ++s)
      if (b->used[s]) fill_diag(b, s, &diag[s]);
  return PF_OK;
}

int pf_batch_launch(pf_batch* b, int nsteps, const int32_t* active) {
  if (b) b->serving = false;
  int rc = arm(b, nsteps, active);
  if (rc) return rc;
  if (*b->hactive <= 0) return PF_OK;
  b->poll_chunk = 16;
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
  if (b->timing) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev_t0, b->ev_t1));
    b->gpu_ms += ms;
  }
  if (b->serving) {
    *done = 1;
    return PF_OK;
  }
  if (*b->hactive <= 0) {
    int rc = pull_scal(b);
    if (rc) return rc;
    *done = 1;
    return PF_OK;
  }
  if (b->rounds > b->round_guard) return fail(PF_ESTATE, "round limit exceeded (internal error)");
  b->poll_chunk = std::min<int64_t>(b->poll_chunk * 2, b->max_chunk);
  int rc = chunk(b, b->poll_chunk);
  if (rc) return rc;
  *done = 0;
  return PF_OK;
}

// Continuous batching: one cavity is armed while the others keep their
// device-resident state; its phase machine starts with the next round.
int pf_batch_start(pf_batch* b, int slot, int nsteps) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (nsteps <= 0) return fail(PF_EINVAL, "nsteps must be > 0");
  const int par = (int)((b->round + 1) & 1);  // buffer of the last round
  HIPCHK(hipMemcpyAsync(b->hscal + slot, b->dscal + (size_t)par * b->max_cav + slot, sizeof(PfScal),
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  PfScal& sc = b->hscal[slot];
  if (sc.phase != PH_IDLE && sc.phase != PH_DONE) return fail(PF_ESTATE, "cavity is already running");
  const PfCav& c = b->hcav[slot];
  if (b->dreset[slot]) {
    sc.dready = 0;
    b->dreset[slot] = 0;
  }
  sc.pending = PH_IDLE;
  sc.done_now = 0;
  sc.steps_done = 0;
  sc.it_run = 0;
  sc.outer_run = 0;
  sc.rounds_run = 0;
  sc.step_round = 0;
  sc.step_len = 0;
  sc.dsetup = 0;
  if (c.mode == PF_MODE_STEP) {
    sc.phase = PH_STATS;
    sc.steps_left = nsteps;
  } else if (c.mode == PF_MODE_PROJECT) {
    sc.phase = PH_PRED_A;
    sc.steps_left = 1;
  } else {
    enter_pinit(c, sc);
    sc.steps_left = 1;
  }
  rc = push_scal(b, slot);
  if (rc) return rc;
  *b->hactive += 1;
  HIPCHK(hipMemcpyAsync(b->dactive, b->hactive, sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->serving = true;
  return PF_OK;
}

int pf_batch_advance(pf_batch* b, int64_t rounds) {
  if (!b || rounds <= 0) return fail(PF_EINVAL, "bad arguments");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  int rc = sync_tables(b);
  if (rc) return rc;
  b->serving = true;
  b->obs_slot = -1;
  if (b->total_blocks <= 0) return PF_OK;
  return chunk(b, rounds);
}

int pf_batch_finished(pf_batch* b, int32_t* finished) {
  if (!b || !finished) return fail(PF_EINVAL, "bad arguments");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  int rc = pull_scal(b);
  if (rc) return rc;
  for (int s = 0; s < b->max_cav; ++s) finished[s] = (b->used[s] && b->hscal[s].phase == PH_DONE) ? 1 : 0;
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
  if (c.ny > b->gather_cap) {
    if (b->dgather) HIPCHK(hipFree(b->dgather));
    if (b->hgather) HIPCHK(hipHostFree(b->hgather));
    b->dgather = nullptr;
    b->hgather = nullptr;
    b->gather_cap = 0;
    HIPCHK(hipMalloc((void**)&b->dgather, sizeof(double) * c.ny));
    HIPCHK(hipHostMalloc((void**)&b->hgather, sizeof(double) * c.ny, hipHostMallocDefault));
    b->gather_cap = c.ny;
  }
  // One gather of u[:, nx//2] into a contiguous buffer and one copy: a
  // strided copy of ny single elements costs milliseconds at 1024 rows.
  const int cx = c.nx / 2;
  const dim3 grid((unsigned)((c.ny + 255) / 256)), blk(256);
  if (b->prec == PF_PREC_MIXED)
    k_column<float><<<grid, blk, 0, b->stream>>>((const float*)c.u, c.nx, c.ny, cx, b->dgather);
  else
    k_column<double><<<grid, blk, 0, b->stream>>>((const double*)c.u, c.nx, c.ny, cx, b->dgather);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->hgather, b->dgather, sizeof(double) * c.ny, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  std::memcpy(out, b->hgather, sizeof(double) * c.ny);
  return PF_OK;
}

int pf_batch_set_graph(pf_batch* b, int rounds) {
  if (!b || rounds < 0) return fail(PF_EINVAL, "bad arguments");
  if (b->chunk_pending) return fail(PF_ESTATE, "batch is running");
  drop_graphs(b);
  b->graph_len = graph_len_of(rounds);
  return PF_OK;
}

int pf_batch_set_timing(pf_batch* b, int enable) {
  if (!b) return fail(PF_EINVAL, "batch is NULL");
  b->timing = enable != 0;
  b->gpu_ms = 0.0;
  b->timed_rounds = 0;
  return PF_OK;
}

int pf_batch_timing(pf_batch* b, double* gpu_ms, int64_t* rounds) {
  if (!b || !gpu_ms || !rounds) return fail(PF_EINVAL, "bad arguments");
  *gpu_ms = b->gpu_ms;
  *rounds = b->timed_rounds;
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

int pf_solve_noalloc(int precision, const pf_cavity_params* params, const double* rhs, double* x, void* stream,
                     pf_step_diag* diag) {
  if (!rhs || !x) return fail(PF_EINVAL, "rhs and x must be device pointers");
  if (precision != PF_PREC_FP64 && precision != PF_PREC_MIXED) return fail(PF_EINVAL, "bad precision");
  int rc = check_params(params);
  if (rc) return rc;
  if (params->mode != PF_MODE_SOLVE) return fail(PF_EINVAL, "pf_solve_noalloc runs PF_MODE_SOLVE only");
  PfCav c{};
  fill_params(c, params, precision);
  c.blk0 = 0;
  // Layout inside the scratch array: descriptor, two state parities, the
  // active counter and the workgroup table, then the cavity's vectors and
  // its partials. The right-hand side and the solution are the caller's:
  // b is only read in PF_MODE_SOLVE, and the FP64 solution (x64, or x in
  // FP64) is written straight into x.
  const size_t n = (size_t)c.nx * c.ny;
  const size_t vb = align256(n * esize(precision));
  const size_t off_cav = 0;
  const size_t off_scal = off_cav + align256(sizeof(PfCav));
  const size_t off_active = off_scal + align256(2 * sizeof(PfScal));
  const size_t off_blk = off_active + 256;
  const size_t off_vec = off_blk + align256(sizeof(int) * (size_t)c.nblk);
  const int nvec = (precision == PF_PREC_MIXED) ? 6 : 5;  // p r r1 d0 d1 (+ x: float32 inner iterate)
  const size_t off_part = off_vec + (size_t)nvec * vb;
  const size_t off_y = off_part + align256((size_t)2 * c.nblk * NPS * sizeof(double));
  const size_t need = off_y + (c.defl ? align256(n * sizeof(float)) : 0);  // y: deflate = 2 only
  if (need > PF_NOALLOC_SCRATCH_BYTES)
    return fail(PF_EINVAL, "a " + std::to_string(c.nx) + "x" + std::to_string(c.ny) + " solve needs " +
                               std::to_string(need) + " bytes of scratch; pf_solve_noalloc holds at most " +
                               std::to_string(PF_NOALLOC_SCRATCH_BYTES) + " (PF_NOALLOC_SCRATCH_BYTES)");
  char* base = nullptr;
  rc = scratch_for_current_device(&base);
  if (rc) return rc;
  if (precision == PF_PREC_MIXED) {
    rc = allow_region_lds();
    if (rc) return rc;
  }
  char* w = base + off_vec;
  c.p = w;
  c.r = w + vb;
  c.r1 = w + 2 * vb;
  c.d0 = w + 3 * vb;
  c.d1 = w + 4 * vb;
  c.u = c.v = c.us = c.vs = c.p;  // not touched by the solve's phases
  c.b = const_cast<double*>(rhs);
  if (precision == PF_PREC_MIXED) {
    c.x = w + 5 * vb;
    c.x64 = x;
  } else {
    c.x = x;
    c.x64 = x;
  }
  c.part = (double*)(base + off_part);
  c.y = c.defl ? (float*)(base + off_y) : nullptr;  // solved again on every call
  PfCav* dcav = (PfCav*)(base + off_cav);
  PfScal* dscal = (PfScal*)(base + off_scal);
  int* dactive = (int*)(base + off_active);
  int* dblk = (int*)(base + off_blk);
  PfScal s0;
  std::memset(&s0, 0, sizeof(s0));
  enter_pinit(c, s0);
  s0.pending = PH_IDLE;
  s0.steps_left = 1;
  s0.dt_prev = 1e-2;
  hipStream_t st = (hipStream_t)stream;
  // Descriptor, states, counter and table are written by a kernel from its
  // arguments: no host staging buffer, no memset.
  k_noalloc_init<<<dim3((unsigned)((c.nblk + 255) / 256)), dim3(256), 0, st>>>(c, s0, dcav, dscal, dactive, dblk,
                                                                              c.nblk);
  HIPCHK(hipGetLastError());
  // The kernel instance a one-cavity batch of this cavity would run.
  pf_batch tb;
  tb.prec = precision;
  tb.total_blocks = c.nblk;
  tb.single_slot = 0;
  tb.max_tpb = c.tpb;
  tb.lds_floats = lds_floats_of(c);
  tb.any_xd = c.xd != 0;
  tb.xd_kinds = c.xd ? (c.ca ? 2 : 1) : 0;
  tb.multi = c.rpw > 1;
  tb.s3 = c.ca == 2;
  tb.occ2 = (precision == PF_PREC_MIXED) && env_switch("PYFLOW_OCC2", 1) != 0 && c.ca == 1 && !c.xd && c.tpb <= 2 &&
            c.rpw == 1;
  tb.unroll = env_int("PYFLOW_UNROLL", 2);
  const RoundLaunch L = round_launch(&tb);
  KArgs K;
  std::memset(&K, 0, sizeof(K));
  K.cavs = dcav;
  K.scal = dscal;
  K.blk_cav = dblk;
  K.active = dactive;
  K.max_cav = 1;
  K.slot0 = 0;
  K.cav0 = c;
  const int64_t guard = (c.defl ? 6 : 4) * (int64_t)std::max(c.maxiter, 0) + (c.defl ? 224 : 144);
  int64_t round = 0, chunk_len = 64;
  int host_active = 1;
  while (true) {
    for (int64_t k = 0; k < chunk_len; ++k, ++round) {
      K.parity = (int)(round & 1);
      void* args[] = {&K};
      HIPCHK(hipLaunchKernel(L.fn, L.grid, L.blk, args, L.lds, st));
    }
    HIPCHK(hipMemcpyAsync(&host_active, dactive, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (host_active <= 0) break;
    if (round > guard) return fail(PF_ESTATE, "round limit exceeded (internal error)");
    chunk_len = std::min<int64_t>(chunk_len * 2, 1024);
  }
  if (diag) {
    PfScal fin;
    HIPCHK(hipMemcpyAsync(&fin, dscal + ((round + 1) & 1), sizeof(PfScal), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    diag_from(fin, diag);
  }
  return PF_OK;
}

}  // extern "C"
