// Seismic x-shift registration (matchSeismicSeries): the displacement has an x component only (the
// y component of the field stays zero); the data term is the mean over the image rows of the 1-D
// normalised cross-correlation of the template row shifted by that row of phi_x against the reference
// row (FELineShiftRegisEnergy, registration.h:2137-2193, over NormalizedCrossCorrelationEnergy on the
// row's 1-D grid, registration.h:1909-2018); the regulariser is lambda/2 phi_x^T (K_xx + y K_yy) phi_x
// with y = yRegFactor (findXShift, matchSeries.h:805-868).
//
// One workgroup per (pair, row): the row's reference and template and the shifted template live in LDS;
// FP64 arithmetic in the reference's operation order per row (assembled 1-D mass matrix of the 2-point
// Gauss rule, FastDeformImage with clamping, normalisation by the mass-weighted mean and deviation).
// The rows' values are summed by the control in a fixed order.
#pragma once

#include "ms_kernels.h"

namespace msg {

constexpr int XS_THREADS = 256;

// Local 1-D mass matrix of one element (2-point rule, weights 1/2) times h: l[0] = L00 = L11, l[1] = L01.
__device__ __forceinline__ void xs_mass_local(double h, double &l00, double &l01, double &l11) {
  double L[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double xi = q ? MSG_QP1 : MSG_QP0;
    const double b[2] = {1.0 - xi, xi};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) L[a][c] += 0.5 * b[a] * b[c];
  }
  l00 = L[0][0] * h;
  l01 = L[0][1] * h;
  l11 = L[1][1] * h;
}

// (M v)_i of the assembled 1-D mass matrix: row i summed in ascending column order; the diagonal as
// assembled (element i-1's L11, then element i's L00).
__device__ __forceinline__ double xs_mass_row(const double *v, int i, int n, double l00, double l01, double l11) {
  double dg = 0;
  if (i > 0) dg += l11;
  if (i < n - 1) dg += l00;
  double s = 0;
  if (i > 0) s += l01 * v[i - 1];
  s += dg * v[i];
  if (i < n - 1) s += l01 * v[i + 1];
  return s;
}

// aol::Vector::interpolate (vec.h:461-471)
__device__ __forceinline__ double xs_interp(const double *v, int n, double pos) {
  const int p = (int)pos;
  const double pf = pos - p;
  if (p == n - 1) return v[p];
  return (1 - pf) * v[p] + pf * v[p + 1];
}

// Block-wide FP64 sum of one value per thread (XS_THREADS threads); every thread gets the result.
__device__ __forceinline__ double xs_sum(double v, double *red) {
  const double s = block_sum<XS_THREADS>(v, red);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = s;
  __syncthreads();
  const double r = bc;
  __syncthreads();
  return r;
}

// The reference row's normalisation: u = r - m_r in ul[], and its deviation v_r (registration.h:1934-1947).
__device__ __forceinline__ void xs_reference(double *rl, int n, double l00, double l01, double l11, double *red,
                                             double &vr) {
  double a = 0;
  for (int i = threadIdx.x; i < n; i += XS_THREADS) a += xs_mass_row(rl, i, n, l00, l01, l11);
  const double mr = xs_sum(a, red);
  for (int i = threadIdx.x; i < n; i += XS_THREADS) rl[i] += -mr;
  __syncthreads();
  double b = 0;
  for (int i = threadIdx.x; i < n; i += XS_THREADS) b += rl[i] * xs_mass_row(rl, i, n, l00, l01, l11);
  vr = sqrt(xs_sum(b, red));
  __syncthreads();
}

// Deformed row def[] (FastDeformImage<1-D>: t at i + y_i / h, clamped to [0, n-1]) -> normalised in
// place; returns the row energy -(M tn) . rn with rn = u / v_r (u in ul[]); var_t receives the deviation.
// Expects def[] written and synchronised.
__device__ __forceinline__ double xs_row_energy(double *def, const double *ul, double vr, int n, double l00,
                                                double l01, double l11, double *red, double &var_t) {
  double a = 0;
  for (int i = threadIdx.x; i < n; i += XS_THREADS) a += xs_mass_row(def, i, n, l00, l01, l11);
  const double mt = xs_sum(a, red);
  for (int i = threadIdx.x; i < n; i += XS_THREADS) def[i] += -mt;
  __syncthreads();
  double b = 0;
  for (int i = threadIdx.x; i < n; i += XS_THREADS) b += def[i] * xs_mass_row(def, i, n, l00, l01, l11);
  const double vt = sqrt(xs_sum(b, red));
  var_t = vt;
  const bool dt = !(fabs(vt - 0.) < 1e-15), dr = !(fabs(vr - 0.) < 1e-15);
  if (dt)
    for (int i = threadIdx.x; i < n; i += XS_THREADS) def[i] /= vt;
  __syncthreads();
  double c = 0;
  for (int i = threadIdx.x; i < n; i += XS_THREADS) {
    const double rn = dr ? ul[i] / vr : ul[i];
    c += xs_mass_row(def, i, n, l00, l01, l11) * rn;
  }
  return -xs_sum(c, red);
}

// K_xx and K_yy local matrices of the Q1 element (2x2 rule), as the reference's FEOpMixedDerivative
// assembles them.
__device__ __forceinline__ void xs_stiff_local(double Kx[4][4], double Ky[4][4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Kx[a][b] = Ky[a][b] = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double qx = (q >> 1) ? MSG_QP1 : MSG_QP0, qy = (q & 1) ? MSG_QP1 : MSG_QP0;
    double gx[4], gy[4];
    gx[0] = -(1 - qy); gx[1] = (1 - qy); gx[2] = -qy; gx[3] = qy;
    gy[0] = -(1 - qx); gy[1] = -qx; gy[2] = (1 - qx); gy[3] = qx;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        Kx[a][b] += gx[a] * gx[b] * 0.25;
        Ky[a][b] += gy[a] * gy[b] * 0.25;
      }
  }
}

// ================================================================================================
// Energy pass: per (pair, row) and candidate step size, the row's NCC energy (term 0) and the
// regulariser of the element row below it (term 1; psi = phi + tau d at the stored positions).
// Dynamic LDS: 3 n doubles.
// ================================================================================================
template <typename S>
__global__ void __launch_bounds__(XS_THREADS)
k_energy_xshift(const S *__restrict__ R, const S *__restrict__ T, const S *__restrict__ phi, const S *__restrict__ z,
                const S *__restrict__ dold, int n, double h, double yreg, const PairState *st, double *__restrict__ part,
                int npairs) {
  const int pair = blockIdx.x / n, line = blockIdx.x - pair * n;
  const int nc = st[pair].ncand;
  if (nc == 0) return;  // uniform across the workgroup
  const int dm = st[pair].dirmode;
  const bool cg = (dm == DIR_CG), use_old = st[pair].cg_old != 0, use_dir = (dm != DIR_NONE);
  const double beta = st[pair].beta;
  extern __shared__ double xs_lds[];
  double *tl = xs_lds, *ul = xs_lds + n, *def = xs_lds + 2 * n;
  __shared__ double red[4];
  const size_t nn = (size_t)n * n;
  const S *Rr = R + pair * nn + (size_t)line * n, *Tr = T + pair * nn + (size_t)line * n;
  const S *px = phi + pair * 2 * nn, *zx = z + pair * 2 * nn, *ox = dold + pair * 2 * nn;
  const bool skip = (Rr[0] != Rr[0]) || (Tr[0] != Tr[0]);  // a row starting with NaN is left out
  double l00, l01, l11;
  xs_mass_local(h, l00, l01, l11);
  double Kx[4][4], Ky[4][4];
  xs_stiff_local(Kx, Ky);
  double vr = 0;
  if (!skip) {
    for (int i = threadIdx.x; i < n; i += XS_THREADS) {
      tl[i] = (double)Tr[i];
      ul[i] = (double)Rr[i];
    }
    __syncthreads();
    xs_reference(ul, n, l00, l01, l11, red, vr);
  }
  const int nblk = n;
  for (int k = 0; k < nc; ++k) {
    const double tau = st[pair].cand[k];
    double e_row = 0;
    if (!skip) {
      for (int i = threadIdx.x; i < n; i += XS_THREADS) {
        const size_t g = (size_t)line * n + i;
        const double f = (double)px[g];
        const double y = use_dir ? (double)step_at<S>(px[g], tau, search_dir<S>(zx, ox, g, cg, use_old, beta)) : f;
        double ds = i + y / h;
        ds = ds < 0 ? 0 : (ds > (double)(n - 1) ? (double)(n - 1) : ds);
        def[i] = xs_interp(tl, n, ds);
      }
      __syncthreads();
      double vt;
      e_row = xs_row_energy(def, ul, vr, n, l00, l01, l11, red, vt);
    }
    // regulariser of the elements between rows line and line + 1
    double reg = 0;
    if (line < n - 1) {
      for (int e = threadIdx.x; e < n - 1; e += XS_THREADS) {
        double p[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const size_t g = (size_t)(line + (b >> 1)) * n + e + (b & 1);
          p[b] = use_dir ? (double)step_at<S>(px[g], tau, search_dir<S>(zx, ox, g, cg, use_old, beta)) : (double)px[g];
        }
        double s = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          double r = 0;
#pragma unroll
          for (int b = 0; b < 4; ++b) r += (Kx[a][b] + yreg * Ky[a][b]) * p[b];
          s += p[a] * r;
        }
        reg += 0.5 * s;
      }
    }
    const double rs = block_sum<XS_THREADS>(reg, red);
    if (threadIdx.x == 0) {
      part[((size_t)(pair * MAX_CAND + k) * 2 + 0) * nblk + line] = e_row;
      part[((size_t)(pair * MAX_CAND + k) * 2 + 1) * nblk + line] = rs;
    }
    __syncthreads();
  }
}

// ================================================================================================
// Variation: DE_x = (row derivative of the NCC energy) / n + lambda (K_xx + y K_yy) phi_x, DE_y = 0.
// Row derivative (VariationOfNormalizedCrossCorrelationEnergy / CrossCorrelationForce on the row):
// w = (rn + E tn) / var_t, and at each quad point whose shifted position stays on the row
// -w(x_q) T'(x_q + phi(x_q)), integrated against the two basis functions. Dynamic LDS: 3 n doubles.
// ================================================================================================
template <typename S>
__global__ void __launch_bounds__(XS_THREADS)
k_grad_xshift(const S *__restrict__ R, const S *__restrict__ T, const S *__restrict__ phi, S *__restrict__ DE, int n,
              double h, double yreg, const PairState *st, double lambda_default, int npairs) {
  const int pair = blockIdx.x / n, line = blockIdx.x - pair * n;
  if (st && !participates(st[pair])) return;  // uniform across the workgroup
  const double lambda = st ? st[pair].lambda : lambda_default;
  extern __shared__ double xs_lds[];
  double *tl = xs_lds, *ul = xs_lds + n, *w = xs_lds + 2 * n;
  __shared__ double red[4];
  const size_t nn = (size_t)n * n;
  const S *Rr = R + pair * nn + (size_t)line * n, *Tr = T + pair * nn + (size_t)line * n;
  const S *px = phi + pair * 2 * nn;
  const S *y = px + (size_t)line * n;
  S *o = DE + pair * 2 * nn;
  const bool skip = (Rr[0] != Rr[0]) || (Tr[0] != Tr[0]);
  double l00, l01, l11;
  xs_mass_local(h, l00, l01, l11);
  if (!skip) {
    for (int i = threadIdx.x; i < n; i += XS_THREADS) {
      tl[i] = (double)Tr[i];
      ul[i] = (double)Rr[i];
    }
    __syncthreads();
    double vr;
    xs_reference(ul, n, l00, l01, l11, red, vr);
    for (int i = threadIdx.x; i < n; i += XS_THREADS) {
      double ds = i + (double)y[i] / h;
      ds = ds < 0 ? 0 : (ds > (double)(n - 1) ? (double)(n - 1) : ds);
      w[i] = xs_interp(tl, n, ds);
    }
    __syncthreads();
    double vt;
    const double e = xs_row_energy(w, ul, vr, n, l00, l01, l11, red, vt);  // w now holds tn
    const bool dt = !(fabs(vt - 0.) < 1e-15), dr = !(fabs(vr - 0.) < 1e-15);
    for (int i = threadIdx.x; i < n; i += XS_THREADS) {
      double v = dr ? ul[i] / vr : ul[i];
      v += e * w[i];
      if (dt) v /= vt;
      w[i] = v;
    }
    __syncthreads();
  }
  double Kx[4][4], Ky[4][4];
  xs_stiff_local(Kx, Ky);
  for (int i = threadIdx.x; i < n; i += XS_THREADS) {
    double d = 0;
    if (!skip) {
      // contributions of element i-1 (its dof 1), then element i (dof 0), as assembled
      for (int el = i - 1; el <= i; ++el) {
        if (el < 0 || el > n - 2) continue;
        const int dof = i - el;
        double a = 0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const double xi = q ? MSG_QP1 : MSG_QP0;
          const double offq = (1 - xi) * (double)y[el] + xi * (double)y[el + 1];
          const double tc = (double)el + xi + offq / h;
          int tel = (short)tc;
          double tloc = tc - tel;
          if (tel == n - 1 && tloc == 0) {
            tel = n - 2;
            tloc = 1;
          }
          double nl = 0;
          if (!(tc < 0. || tel < 0 || tel >= n - 1)) {
            const double g = (tl[tel + 1] - tl[tel]) / h;
            const double wq = (1 - xi) * w[el] + xi * w[el + 1];
            nl = g * -wq;
          }
          a += (dof ? xi : 1 - xi) * 0.5 * nl;
        }
        d += h * a;
      }
      d /= n;
    }
    // (K_xx + y K_yy) phi_x at node (i, line): entries assembled per element (K_xx of the 4 elements,
    // then y K_yy), each row summed over its 3 x 3 neighbours
    double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int pass = 0; pass < 2; ++pass)
      for (int ey = line - 1; ey <= line; ++ey)
        for (int ex = i - 1; ex <= i; ++ex) {
          if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
          const int bi = (i - ex) + 2 * (line - ey);
#pragma unroll
          for (int bj = 0; bj < 4; ++bj) {
            const int jx = ex + (bj & 1), jy = ey + ((bj & 2) >> 1);
            c[jy - line + 1][jx - i + 1] += pass ? yreg * Ky[bi][bj] : Kx[bi][bj];
          }
        }
    double s = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int jx = i + dx, jy = line + dy;
        if (jx < 0 || jy < 0 || jx >= n || jy >= n) continue;
        s += c[dy + 1][dx + 1] * (double)px[(size_t)jy * n + jx];
      }
    double v = 0;
    v += d;
    v += lambda * s;
    o[(size_t)line * n + i] = (S)v;
    o[nn + (size_t)line * n + i] = (S)0;
  }
}

}  // namespace msg
