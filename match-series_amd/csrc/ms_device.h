// Device-side building blocks of the gfx950 registration kernels.
//
// Layout: a level-l image is n*n values (n = 2^l), node (x, y) at x + y*n. A displacement field is
// two such arrays (x component first) in domain units (pixel offset = value / h, h = 1/(n-1)).
// Fields are stored as S (float or double); every kernel computes in double in registers/LDS,
// which on CDNA4 runs at the same vector rate as unpacked FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msg {

// 2-point Gauss rule on [0,1]; quad point q = (QP[q/2], QP[q%2]), weight 1/4
// (baseFunctionSet.h:1535-1554, 1436-1467).
#define MSG_QP0 0.2113248654051871
#define MSG_QP1 0.7886751345948129

__device__ __forceinline__ double qpx(int q) { return (q >> 1) ? MSG_QP1 : MSG_QP0; }
__device__ __forceinline__ double qpy(int q) { return (q & 1) ? MSG_QP1 : MSG_QP0; }

// Q1 basis on the unit square, local dofs (0,0),(1,0),(0,1),(1,1) (baseFunctionSet.h:127-193).
__device__ __forceinline__ double q1(int b, double x, double y) {
  switch (b) {
    case 0: return (1.0 - x) * (1.0 - y);
    case 1: return x * (1.0 - y);
    case 2: return (1.0 - x) * y;
    default: return x * y;
  }
}
__device__ __forceinline__ double q1dx(int b, double y) {
  switch (b) {
    case 0: return y - 1.0;
    case 1: return 1.0 - y;
    case 2: return -y;
    default: return y;
  }
}
__device__ __forceinline__ double q1dy(int b, double x) {
  switch (b) {
    case 0: return x - 1.0;
    case 1: return -x;
    case 2: return 1.0 - x;
    default: return x;
  }
}
// basis value of dof b at quad point q
__device__ __forceinline__ double bq(int b, int q) { return q1(b, qpx(q), qpy(q)); }

__device__ __forceinline__ int l2g(int ex, int ey, int b, int n) {
  return (ex + (b & 1)) + (ey + ((b >> 1) & 1)) * n;
}

template <typename S>
__device__ __forceinline__ double ld(const S *__restrict__ p, int i) {
  return static_cast<double>(p[i]);
}
// round a double to the storage type and back (what a store + reload would give)
template <typename S>
__device__ __forceinline__ double rnd(double v) {
  return static_cast<double>(static_cast<S>(v));
}

// phi + tau * d as stored in S: FP64 fields in FP64 arithmetic, FP32 fields by one fused multiply-add
// (the step sizes are powers of two, exact in FP32). The update kernel, the FP64-arithmetic energy
// passes (k_energy and its FP64 re-evaluation of FP32 pairs, k_energy_exact_f32), the NCC,
// consistency and x-shift passes use this one definition, so a candidate they evaluate is bitwise
// the position the update stores. The FP32-arithmetic SSD pass (k_energy_f32) forms the quad points
// of phi + tau d linearly in tau from one interpolation of phi and d; they differ from those of the
// stored position by FP32 rounding, and a step decision whose margin lies within that pass's error
// bound is taken again on the re-evaluation at the stored positions.
template <typename S>
__device__ __forceinline__ S step_at(S f, double tau, S d);
template <>
__device__ __forceinline__ double step_at<double>(double f, double tau, double d) {
  return f + tau * d;
}
template <>
__device__ __forceinline__ float step_at<float>(float f, double tau, float d) {
  return __fmaf_rn((float)tau, d, f);
}

// Deformed quad point: transformAndClipCoord (deformations.h:795-821) followed by
// getLocalCoordsRegularRectangularGrid (configurators.h:32-50), in domain units as the reference.
struct Hit {
  int ex, ey;
  double lx, ly;
  bool inx, iny;
};

__device__ __forceinline__ void clip_axis(double coord, int n, double h, int width, int &el, double &loc,
                                          bool &within) {
  within = true;
  if (coord < 0.) {
    coord = 0.;
    within = false;
  } else if (coord >= width) {
    coord = width - 2.220446049250313e-16;  // DBL_EPSILON
    within = false;
  }
  const double sc = coord / h;
  el = static_cast<int>(sc);
  loc = sc - el;
  if (el == n - 1 && fabs(loc - 0.0) < 1e-15) {
    el = n - 2;
    loc = 1;
  }
  // Never taken for dyadic grids (the snap above covers every clipped coordinate); keeps the gather
  // inside the image whatever the input.
  if (el < 0) el = 0;
  if (el > n - 2) el = n - 2;
}

__device__ __forceinline__ Hit deform_qp(int ex, int ey, int q, double ox, double oy, int n, double h, int width) {
  Hit r;
  double cx = ((double)ex + qpx(q)) * h;
  double cy = ((double)ey + qpy(q)) * h;
  cx += ox;
  cy += oy;
  clip_axis(cx, n, h, width, r.ex, r.lx, r.inx);
  clip_axis(cy, n, h, width, r.ey, r.ly, r.iny);
  return r;
}

// transformCoord<ConfType, false> (deformations.h:623-668, 681-693): the quad point in grid
// coordinates plus offset/h, without clipping; false when it leaves the grid on either axis (the
// reference stops at the first such axis; only the flag matters then). The element / local
// coordinate of a point exactly on the last node line is moved into the last element.
__device__ __forceinline__ bool transform_noclip(int ex, int ey, int q, double ox, double oy, int n, double h,
                                                 Hit &r) {
  const double c[2] = {(double)ex + qpx(q), (double)ey + qpy(q)};
  const double off[2] = {ox, oy};
  int el[2];
  double loc[2];
  for (int i = 0; i < 2; ++i) {
    const double tc = c[i] + off[i] / h;
    if (!(tc >= 0.)) return false;  // also NaN
    if (tc >= (double)(n - 1)) {
      if (tc != (double)(n - 1)) return false;
      el[i] = n - 2;
      loc[i] = 1;
    } else {
      el[i] = static_cast<int>(tc);
      loc[i] = tc - el[i];
    }
  }
  r.ex = el[0];
  r.ey = el[1];
  r.lx = loc[0];
  r.ly = loc[1];
  r.inx = r.iny = true;
  return true;
}

// T(el, loc): DiscreteFunctionDefault::evaluate (discreteFunction.h:229-236)
template <typename S>
__device__ __forceinline__ double eval_img(const S *__restrict__ T, const Hit &p, int n, double t[4]) {
  const int i0 = p.ex + p.ey * n;
  t[0] = ld(T, i0);
  t[1] = ld(T, i0 + 1);
  t[2] = ld(T, i0 + n);
  t[3] = ld(T, i0 + n + 1);
  double w = 0.;
  for (int b = 0; b < 4; ++b) w += t[b] * q1(b, p.lx, p.ly);
  return w;
}

// ScalarArray<QC_2D>::interpolate (scalarArray.h:499-529) including the Y == 0 branch, on a w x h
// array (w, h >= 2).
template <typename S>
__device__ __forceinline__ double interp_wh(const S *__restrict__ img, int w, int h, double X, double Y) {
  if (X != X) X = 0.;  // NaN coordinates (undefined in the reference) read inside the image
  if (Y != Y) Y = 0.;
  int xL = (int)X, yL = (int)Y;
  if (X >= (double)(w - 1)) xL = w - 2;
  if (Y >= (double)(h - 1)) yL = h - 2;
  if (X < 0.) xL = 0;
  if (Y < 0.) yL = 0;
  X -= xL;
  Y -= yL;
  const int xU = xL + 1, yU = yL + 1;
  const double v0 = ld(img, xL + yL * w), v1 = ld(img, xU + yL * w);
  const double v2 = ld(img, xL + yU * w), v3 = ld(img, xU + yU * w);
  if (Y == 0) return (1 - X) * v0 + X * v1;
  return (1 - X) * (1 - Y) * v0 + X * (1 - Y) * v1 + (1 - X) * Y * v2 + X * Y * v3;
}

template <typename S>
__device__ __forceinline__ double interp(const S *__restrict__ img, int n, double X, double Y) {
  return interp_wh(img, n, n, X, Y);
}


// Nodal warp of FastDeformImage (deformations.h:140-162): T at node (x, y) displaced by (px, py)
// in domain units, clamped to the grid (the second test also sends NaN to 0, keeping the gather
// inside the image).
template <typename S>
__device__ __forceinline__ double fast_warp(const S *__restrict__ T, int n, double h, int x, int y, double px,
                                            double py) {
  const double hi = (double)(n - 1);
  double X = x + px / h, Y = y + py / h;
  X = X > hi ? hi : X;
  X = X >= 0. ? X : 0.;
  Y = Y > hi ? hi : Y;
  Y = Y >= 0. ? Y : 0.;
  return interp(T, n, X, Y);
}

// Q1 element mass matrix times the element volume (FELinScalarWeightedMassInterface::prepareLocalMatrix,
// FEOpInterface.h:1446-1475): sum over the quad points of (b_i b_j * 1) * w_q, then * vol.
__device__ __forceinline__ void mass_local(double vol, double ml[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) ml[i][j] = 0.;
  for (int q = 0; q < 4; ++q)
    for (int i = 0; i < 4; ++i) {
      const double bi = bq(i, q);
      for (int j = i; j < 4; ++j) ml[i][j] += bi * bq(j, q) * 1. * 0.25;
    }
  for (int i = 0; i < 4; ++i)
    for (int j = i; j < 4; ++j) ml[j][i] = ml[i][j] = ml[i][j] * vol;
}

// Element quadratic form psi^T K_loc psi of the Q1 stiffness matrix written in differences:
// K_loc = (1/6)[4 -1 -1 -2; -1 4 -2 -1; -1 -2 4 -1; -2 -1 -1 4] (zero row sums), so
// psi^T K_loc psi = 1/6 (sum over the 4 edges of diff^2) + 2/6 (sum over the 2 diagonals of diff^2).
__device__ __forceinline__ double stiff_form(const double p[4]) {
  const double e0 = p[0] - p[1], e1 = p[0] - p[2], e2 = p[1] - p[3], e3 = p[2] - p[3];
  const double d0 = p[0] - p[3], d1 = p[1] - p[2];
  return (e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3) * (1.0 / 6.0) + (d0 * d0 + d1 * d1) * (2.0 / 6.0);
}

// ------------------------------------------------------------------------------------------------
// Reductions (wave64 shuffles, then LDS across the block's waves). Deterministic: fixed shapes.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over a block of BLOCK threads; the result is valid in thread 0. red must hold BLOCK/64 doubles.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += red[w];
  }
  return s;
}

// Sums of NV values over a block of BLOCK threads in one pass (valid in thread 0); red holds
// (BLOCK / 64) * NV doubles. Fixed order.
template <int NV, int BLOCK = 256>
__device__ __forceinline__ void block_sum_n(double (&v)[NV], double *red) {
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = red[k];
      for (int w = 1; w < BLOCK / 64; ++w) a += red[w * NV + k];
      v[k] = a;
    }
}

// Map a linear workgroup id so that each of the 8 XCDs receives a contiguous range of tiles
// (workgroups are dispatched round-robin over XCDs); neighbouring tiles then share an L2.
__device__ __forceinline__ int xcd_remap(int wg, int total) {
  const int nx = 8;
  const int xcd = wg % nx, local = wg / nx;
  const int base = total / nx, extra = total % nx;
  return xcd * base + (xcd < extra ? xcd : extra) + local;
}

// ------------------------------------------------------------------------------------------------
// FD heat operator of the H1 metric (HeatEquation2DFD::apply, linearSmoothOp.cpp:59-146):
// A u = u + hs (4u - sum of the 4 neighbours), neighbours outside the grid mirrored inward
// (which doubles the inward neighbour on edges and corners). get(x, y) reads u in-domain.
// ------------------------------------------------------------------------------------------------
template <typename Get>
__device__ __forceinline__ double heat_apply(const Get &get, int x, int y, int n, double hs) {
  if (n == 1) return get(0, 0);
  const double u = get(x, y);
  const double a = u * 4;
  const int e = n - 1;
  const bool bx0 = (x == 0), bx1 = (x == e), by0 = (y == 0), by1 = (y == e);
  if (!(bx0 || bx1 || by0 || by1)) {
    double b = 0.0;
    b -= get(x + 1, y);
    b -= get(x - 1, y);
    b -= get(x, y + 1);
    b -= get(x, y - 1);
    return u + hs * (a + b);
  }
  if ((bx0 || bx1) && (by0 || by1)) {  // corner
    const int ox = bx0 ? 1 : -1, oy = by0 ? 1 : -1;
    const double b = -get(x + ox, y) - get(x, y + oy);
    return u + hs * (a + 2 * b);
  }
  if (by0 || by1) {  // edge parallel to the x axis
    const int oy = by0 ? 1 : -1;
    double b = -2 * (get(x, y + oy));
    b -= get(x + 1, y);
    b -= get(x - 1, y);
    return u + hs * (a + b);
  }
  const int ox = bx0 ? 1 : -1;  // edge parallel to the y axis
  double b = -2 * (get(x + ox, y));
  b -= get(x, y + 1);
  b -= get(x, y - 1);
  return u + hs * (a + b);
}

}  // namespace msg
