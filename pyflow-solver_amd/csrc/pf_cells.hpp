// Cell-level discrete operators of PyFlow's projection step.
//
// Every kernel (standalone operators, the fused step phases, the batched
// rounds) evaluates cells through these functions, so all of them produce the
// same bits for the same inputs. With T = double and the library compiled with
// -ffp-contract=off, each function reproduces the reference's float64 NumPy
// expression element for element (same operands, same order, divisions kept
// as divisions). Accessors are callables f(j, i) -> T over an (ny, nx) grid;
// callers only ask for in-range cells.
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

// Python's min/max on two floats: min(a, b) keeps a unless b < a.
__device__ __host__ inline double pymin(double a, double b) { return (b < a) ? b : a; }
__device__ __host__ inline double pymax(double a, double b) { return (b > a) ? b : a; }

// np.max-style maximum: any NaN operand makes the result NaN.
__device__ inline double nanmax(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return (b > a) ? b : a;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Coherent 5-point Laplacian = div_c(grad_c(f)) at (j,i), from the centre value
// and its in-range neighbours (core_ops_coherent.py:45-84):
//   gx(j,i) = (f[j,i+1]-f[j,i])/dx, 0 at i = nx-1;   gy likewise along y;
//   lap = (0 + xpart) + ypart,  xpart = (gx[i]-gx[i-1])/dx (i>0) or gx[0]/dx.
// Degenerate grids (nx<2 or ny<2) give 0.
template <class T>
__device__ inline T lap_coherent5(T c, T e, T w, T n, T s, int nx, int ny, int j, int i, T dx, T dy) {
  if (nx < 2 || ny < 2) return T(0);
  const T gxc = (i < nx - 1) ? (e - c) / dx : T(0);
  const T gyc = (j < ny - 1) ? (n - c) / dy : T(0);
  T xp, yp;
  if (i >= 1) {
    const T gxw = (c - w) / dx;
    xp = (gxc - gxw) / dx;
  } else {
    xp = gxc / dx;
  }
  if (j >= 1) {
    const T gys = (c - s) / dy;
    yp = (gyc - gys) / dy;
  } else {
    yp = gyc / dy;
  }
  return (T(0) + xp) + yp;
}

// Pressure operators (pressure_solver.py:97-105 and :69-77).
//   op 0: -lap, row 0 replaced by the identity (pinned reference cell)
//   op 1: +lap, every boundary row replaced by the identity (Dirichlet)
template <class T>
__device__ inline T apply_op5(int op, T c, T e, T w, T n, T s, int nx, int ny, int j, int i, T dx, T dy) {
  const T lap = lap_coherent5<T>(c, e, w, n, s, nx, ny, j, i, dx, dy);
  if (op == 0) {
    if (j == 0 && i == 0) return c;
    return -lap;
  }
  if (j == 0 || j == ny - 1 || i == 0 || i == nx - 1) return c;
  return lap;
}

// Float32 form of the same operators for the mixed-precision path: the
// reflective-Neumann 5-point stencil ((e-c)-(c-w))*cx + ((n-c)-(c-s))*cy with
// cx = 1/dx^2, cy = 1/dy^2 precomputed, missing neighbours' differences
// dropped. Algebraically identical to lap_coherent5; rounds differently.
__device__ inline float lap5_fast(float c, float e, float w, float n, float s, int nx, int ny, int j, int i,
                                  float cx, float cy) {
  if (nx < 2 || ny < 2) return 0.0f;
  const float ex = (i < nx - 1) ? (e - c) : 0.0f;
  const float wx = (i > 0) ? (c - w) : 0.0f;
  const float ny_ = (j < ny - 1) ? (n - c) : 0.0f;
  const float sy = (j > 0) ? (c - s) : 0.0f;
  return __fmul_rn(ex - wx, cx) + __fmul_rn(ny_ - sy, cy);
}
__device__ inline float apply_op5_fast(int op, float c, float e, float w, float n, float s, int nx, int ny, int j,
                                       int i, float cx, float cy) {
  const float lap = lap5_fast(c, e, w, n, s, nx, ny, j, i, cx, cy);
  if (op == 0) {
    if (j == 0 && i == 0) return c;
    return -lap;
  }
  if (j == 0 || j == ny - 1 || i == 0 || i == nx - 1) return c;
  return lap;
}

// Gather the 5-point neighbourhood through an accessor and apply the operator.
template <class T, class F>
__device__ inline T apply_op_at(int op, F f, int nx, int ny, int j, int i, T dx, T dy) {
  const T c = f(j, i);
  const T e = (i + 1 < nx) ? f(j, i + 1) : T(0);
  const T w = (i > 0) ? f(j, i - 1) : T(0);
  const T n = (j + 1 < ny) ? f(j + 1, i) : T(0);
  const T s = (j > 0) ? f(j - 1, i) : T(0);
  return apply_op5<T>(op, c, e, w, n, s, nx, ny, j, i, dx, dy);
}

// Coherent gradient components (core_ops_coherent.py:45-61).
template <class T, class F>
__device__ inline T grad_x_coherent(F f, int nx, int ny, int j, int i, T dx) {
  if (nx < 2 || ny < 2 || i >= nx - 1) return T(0);
  return (f(j, i + 1) - f(j, i)) / dx;
}
template <class T, class F>
__device__ inline T grad_y_coherent(F f, int nx, int ny, int j, int i, T dy) {
  if (nx < 2 || ny < 2 || j >= ny - 1) return T(0);
  return (f(j + 1, i) - f(j, i)) / dy;
}

// Coherent (backward) divergence (core_ops_coherent.py:63-79).
template <class T, class FU, class FV>
__device__ inline T div_coherent(FU U, FV V, int nx, int ny, int j, int i, T dx, T dy) {
  if (nx < 2 || ny < 2) return T(0);
  const T xp = (i >= 1) ? (U(j, i) - U(j, i - 1)) / dx : U(j, 0) / dx;
  const T yp = (j >= 1) ? (V(j, i) - V(j - 1, i)) / dy : V(0, i) / dy;
  return (T(0) + xp) + yp;
}

// Central divergence with one-sided edges (fluid_ops.py:8-28). The reference
// assigns interior, left column, row 0, row ny-1 and the four corners; the
// right column between the corners is never written and stays 0.
// Requires nx >= 2 and ny >= 2.
template <class T, class FU, class FV>
__device__ inline T div_central(FU U, FV V, int nx, int ny, int j, int i, T dx, T dy) {
  const bool jlo = (j == 0), jhi = (j == ny - 1), ilo = (i == 0), ihi = (i == nx - 1);
  const T two_dx = T(2) * dx, two_dy = T(2) * dy;
  if ((jlo || jhi) && (ilo || ihi)) {
    const T xt = ilo ? (U(j, 1) - U(j, 0)) / dx : (U(j, nx - 1) - U(j, nx - 2)) / dx;
    const T yt = jlo ? (V(1, i) - V(0, i)) / dy : (V(ny - 1, i) - V(ny - 2, i)) / dy;
    return xt + yt;
  }
  if (ihi) return T(0);
  if (ilo) return (U(j, 1) - U(j, 0)) / dx + (V(j + 1, 0) - V(j - 1, 0)) / two_dy;
  if (jlo) return (U(0, i + 1) - U(0, i - 1)) / two_dx + (V(1, i) - V(0, i)) / dy;
  if (jhi) return (U(j, i + 1) - U(j, i - 1)) / two_dx + (V(j, i) - V(j - 1, i)) / dy;
  return (U(j, i + 1) - U(j, i - 1)) / two_dx + (V(j + 1, i) - V(j - 1, i)) / two_dy;
}

// QUICK face value ((3/8) a + (6/8) b) - (1/8) c, left to right.
template <class T>
__device__ inline T quick_face(T a, T b, T c) {
  return (T(0.375) * a + T(0.75) * b) - T(0.125) * c;
}

// QUICK convective term (advection.py:45-84): edge-padded stencil (indices
// clamped), cell-centred velocity on both faces, branch on >= 0, and the
// reference's negative-flow faces reproduced as written.
template <class T, class F>
__device__ inline T adv_quick(T uc, T vc, F f, int nx, int ny, int j, int i, T dx, T dy) {
  const T P = f(j, i);
  const T E = f(j, clampi(i + 1, 0, nx - 1));
  const T W = f(j, clampi(i - 1, 0, nx - 1));
  const T WW = f(j, clampi(i - 2, 0, nx - 1));
  const T N = f(clampi(j + 1, 0, ny - 1), i);
  const T S = f(clampi(j - 1, 0, ny - 1), i);
  const T SS = f(clampi(j - 2, 0, ny - 1), i);
  T pe, pw, pn, ps;
  if (uc >= T(0)) {
    pe = quick_face<T>(E, P, W);
    pw = quick_face<T>(P, W, WW);
  } else {
    pe = quick_face<T>(W, P, E);
    pw = quick_face<T>(WW, W, P);
  }
  if (vc >= T(0)) {
    pn = quick_face<T>(N, P, S);
    ps = quick_face<T>(P, S, SS);
  } else {
    pn = quick_face<T>(S, P, N);
    ps = quick_face<T>(SS, S, P);
  }
  return (uc * pe - uc * pw) / dx + (vc * pn - vc * ps) / dy;
}

// First-order upwind term (advection.py:17-43): branch on > 0; (f-f)/d at the
// low edge; forward neighbour clamped at the high edge.
template <class T, class F>
__device__ inline T adv_upwind(T uc, T vc, F f, int nx, int ny, int j, int i, T dx, T dy) {
  const T P = f(j, i);
  T fx, fy;
  if (uc > T(0)) {
    fx = (i > 0) ? (P - f(j, i - 1)) / dx : (P - P) / dx;
  } else {
    fx = (f(j, (i + 1 < nx) ? i + 1 : i) - P) / dx;
  }
  if (vc > T(0)) {
    fy = (j > 0) ? (P - f(j - 1, i)) / dy : (P - P) / dy;
  } else {
    fy = (f((j + 1 < ny) ? j + 1 : j, i) - P) / dy;
  }
  return uc * fx + vc * fy;
}

// Jacobi inverse diagonal (pressure_solver.py:15-40, preconditioners.py:18-21):
// 1 on every boundary cell, 1/(2/dx^2+2/dy^2) inside.
__device__ inline double jacobi_minv(int jacobi, double invd, int nx, int ny, int j, int i) {
  if (!jacobi) return 1.0;
  const bool bnd = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
  return bnd ? 1.0 : invd;
}

}  // namespace pf
