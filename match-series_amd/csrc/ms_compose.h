// Deformation composition with harmonic extension: ConcatAndSmoothlyExtendDeformations
// (quoc/deformations.h:499-520) and SmoothlyExtendImage (L364-404). The extension solves the
// Q1 Laplace problem with Dirichlet data on the defined nodes by diagonally preconditioned CG
// (PCGInverse, aol/solver.h:264-359; absolute stop on |g|^2 < 1e-16, at most 1000 steps), both
// components in the same launches. All scalar work stays on the device; every kernel of an
// iteration reads the component's `done` flag, so iterations launched past convergence change
// nothing and the host only polls the flags every few iterations.
#pragma once

#include "ms_device.h"

namespace msg {

struct PcgScalars {
  double an;     // alpha numerator = <g, h> (also beta denominator)
  double alpha;
  double beta;
  double spn;    // |g|^2
  int it;
  int done;
  int pad[2];
};

// r_c = phi1_c(x + phi2(x)) + phi2_c(x) (NaN where x + phi2(x) leaves the grid), the defined mask,
// and the Dirichlet data bc_c = defined ? r_c : 0.
template <typename S>
__global__ void k_compose_init(const S *__restrict__ phi1, const S *__restrict__ phi2, double *__restrict__ bc,
                               unsigned char *__restrict__ defined, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = n * n;
  if (i >= nn) return;
  const int x = i % n, y = i / n;
  const double h = 1.0 / (double)(n - 1);
  const double dx = x + ld(phi2, i) / h, dy = y + ld(phi2, nn + i) / h;
  const double hi = (double)(n - 1);
  const bool inside = !(dx < 0.0 || dx > hi) && !(dy < 0.0 || dy > hi);
  double r0 = __builtin_nan(""), r1 = __builtin_nan("");
  if (inside) {
    r0 = interp(phi1, n, dx, dy);
    r1 = interp(phi1 + nn, n, dx, dy);
  }
  r0 += ld(phi2, i);
  r1 += ld(phi2, nn + i);
  const bool def = !isnan(r0);
  defined[i] = def ? 1 : 0;
  bc[i] = def ? r0 : 0.;
  bc[nn + i] = def ? r1 : 0.;
}

// Assembled Q1 stiffness row of node (x, y) applied to u (element contributions, K_loc of the
// unit element). masked: only undefined nodes contribute (Dirichlet columns removed).
__device__ __forceinline__ double stiff_row(const double *__restrict__ u, const unsigned char *__restrict__ def,
                                            int x, int y, int n, bool masked) {
  const double kl[4][4] = {{4. / 6, -1. / 6, -1. / 6, -2. / 6},
                           {-1. / 6, 4. / 6, -2. / 6, -1. / 6},
                           {-1. / 6, -2. / 6, 4. / 6, -1. / 6},
                           {-2. / 6, -1. / 6, -1. / 6, 4. / 6}};
  double s = 0;
  for (int ey = y - 1; ey <= y; ++ey)
    for (int ex = x - 1; ex <= x; ++ex) {
      if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
      const int bi = (x - ex) + 2 * (y - ey);
      for (int bj = 0; bj < 4; ++bj) {
        const int j = l2g(ex, ey, bj, n);
        if (masked && def[j]) continue;
        s += kl[bi][bj] * u[j];
      }
    }
  return s;
}

__device__ __forceinline__ double stiff_diag(int x, int y, int n) {
  int ne = 0;
  for (int ey = y - 1; ey <= y; ++ey)
    for (int ex = x - 1; ex <= x; ++ex)
      if (ex >= 0 && ey >= 0 && ex < n - 1 && ey < n - 1) ++ne;
  return ne * (4.0 / 6.0);
}

// rhs = -K bc, overwritten by bc on defined nodes; invd = 1/|diag| of the Dirichlet matrix.
__global__ void k_compose_rhs(const double *__restrict__ bc, const unsigned char *__restrict__ def,
                              double *__restrict__ rhs, double *__restrict__ invd, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = n * n;
  if (i >= 2 * nn) return;
  const int c = i / nn, j = i - c * nn;
  const int x = j % n, y = j / n;
  double v = -stiff_row(bc + c * nn, def, x, y, n, false);
  if (def[j]) v = bc[c * nn + j];
  rhs[i] = v;
  if (c == 0) invd[j] = 1.0 / fabs(def[j] ? 1.0 : stiff_diag(x, y, n));
}

// h = A x with A the Dirichlet stiffness matrix (identity rows on defined nodes).
__device__ __forceinline__ double dir_apply(const double *__restrict__ u, const unsigned char *__restrict__ def, int j,
                                            int n) {
  if (def[j]) return u[j];
  return stiff_row(u, def, j % n, j / n, n, true);
}

template <typename S>
__global__ void k_pcg_start(const S *__restrict__ dest_in, double *__restrict__ x, const double *__restrict__ rhs,
                            const unsigned char *__restrict__ def, const double *__restrict__ invd,
                            double *__restrict__ g, double *__restrict__ hbuf, double *__restrict__ d,
                            double *__restrict__ part, int n, int nblk) {
  __shared__ double red[4];
  const int nn = n * n;
  const int c = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double ss = 0;
  if (j < nn) {
    // x has been filled with the start value by the caller
    const double hv = dir_apply(x + c * nn, def, j, n);
    const double gv = hv - rhs[c * nn + j];
    const double pv = invd[j] * gv;
    g[c * nn + j] = gv;
    hbuf[c * nn + j] = pv;
    d[c * nn + j] = 0.0 - pv;
    ss = gv * gv;
  }
  const double s = block_sum<256>(ss, red);
  if (threadIdx.x == 0) part[c * nblk + blockIdx.x] = s;
}

// h = A d and <d, h>
__global__ void k_pcg_apply(const double *__restrict__ d, const unsigned char *__restrict__ def,
                            double *__restrict__ hbuf, double *__restrict__ part, const PcgScalars *__restrict__ sc,
                            int n, int nblk) {
  const int c = blockIdx.y;
  if (sc[c].done) return;
  __shared__ double red[4];
  const int nn = n * n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double ss = 0;
  if (j < nn) {
    const double hv = dir_apply(d + c * nn, def, j, n);
    hbuf[c * nn + j] = hv;
    ss = d[c * nn + j] * hv;
  }
  const double s = block_sum<256>(ss, red);
  if (threadIdx.x == 0) part[c * nblk + blockIdx.x] = s;
}

// x += alpha d; g += alpha h; h = invd g; partial <g, h> and <g, g>
__global__ void k_pcg_update(double *__restrict__ x, const double *__restrict__ d, double *__restrict__ g,
                             double *__restrict__ hbuf, const double *__restrict__ invd, double *__restrict__ part,
                             const PcgScalars *__restrict__ sc, int n, int nblk) {
  const int c = blockIdx.y;
  if (sc[c].done) return;
  __shared__ double red[4];
  const int nn = n * n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = sc[c].alpha;
  double gh = 0, gg = 0;
  if (j < nn) {
    const int k = c * nn + j;
    x[k] += alpha * d[k];
    const double gv = g[k] + alpha * hbuf[k];
    g[k] = gv;
    const double pv = invd[j] * gv;
    hbuf[k] = pv;
    gh = gv * pv;
    gg = gv * gv;
  }
  const double a = block_sum<256>(gh, red);
  const double b = block_sum<256>(gg, red);
  if (threadIdx.x == 0) {
    part[(2 * c) * nblk + blockIdx.x] = a;
    part[(2 * c + 1) * nblk + blockIdx.x] = b;
  }
}

// d = beta d - h
__global__ void k_pcg_dir(double *__restrict__ d, const double *__restrict__ hbuf, const PcgScalars *__restrict__ sc,
                          int n) {
  const int c = blockIdx.y;
  if (sc[c].done) return;
  const int nn = n * n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nn) {
    const int k = c * nn + j;
    double v = d[k] * sc[c].beta;
    v -= hbuf[k];
    d[k] = v;
  }
}

// Scalar steps, one workgroup per component. stage 0: start (spn, an from <g,h>), 1: alpha from
// <d, A d>, 2: beta and the stopping test.
__global__ void k_pcg_scalar(PcgScalars *__restrict__ sc, const double *__restrict__ part, int nblk, int stage,
                             const double *__restrict__ part2) {
  const int c = blockIdx.x;
  __shared__ double red[4];
  PcgScalars &s = sc[c];
  if (stage != 0 && s.done) return;
  auto sum = [&](const double *p) {
    double v = 0;
    for (int i = threadIdx.x; i < nblk; i += 256) v += p[i];
    return block_sum<256>(v, red);
  };
  if (stage == 0) {
    const double spn = sum(part + c * nblk);      // <g, g>
    const double an = sum(part2 + c * nblk);      // <g, h>
    if (threadIdx.x == 0) {
      s.spn = spn;
      s.an = an;
      s.it = 0;
      s.done = (spn < 1e-16) || isnan(spn) ? 1 : 0;
    }
  } else if (stage == 1) {
    const double ad = sum(part + c * nblk);
    if (threadIdx.x == 0) s.alpha = s.an / ad;
  } else {
    const double bn = sum(part + (2 * c) * nblk);
    const double spn = sum(part + (2 * c + 1) * nblk);
    if (threadIdx.x == 0) {
      s.beta = bn / s.an;
      s.an = bn;
      s.spn = spn;
      s.it += 1;
      s.done = ((spn < 1e-16) || (s.it >= 1000) || isnan(spn)) ? 1 : 0;
    }
  }
}

// <g, h> of the start residual (an of the first step)
__global__ void k_dot_blocks(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ part,
                             int n, int nblk) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  const int nn = n * n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const double v = (j < nn) ? a[c * nn + j] * b[c * nn + j] : 0.0;
  const double s = block_sum<256>(v, red);
  if (threadIdx.x == 0) part[c * nblk + blockIdx.x] = s;
}

template <typename D>
__global__ void k_store(const double *__restrict__ x, D *__restrict__ out, int cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = static_cast<D>(x[i]);
}

}  // namespace msg
