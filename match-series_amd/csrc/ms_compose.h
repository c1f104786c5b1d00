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

// ------------------------------------------------------------------------------------------------
// PCG on the undefined nodes only. The Dirichlet matrix is block diagonal: identity rows on the
// defined nodes D, the masked stiffness K_UU on the undefined nodes U, and the preconditioner is
// the identity on D. On D every CG vector is therefore a scalar multiple of the start residual
// g0 = x0 - bc: g = gam g0, P g = eta g0, d = del g0, x = x0 + xi g0, and its share of every inner
// product is that product of scalars times G0 = |g0|^2. The kernels run over U (a compacted index
// list) and carry the D block as these scalars, so an iteration costs O(|U|) instead of O(n^2).
// The recurrences are the reference's (the oracle's pcgDirichlet) term by term; only the rounding
// of the D block differs (one scaling instead of per-node updates).
//
// Two kernels per iteration k: A forms d_k = beta d_{k-1} - P g (beta from B's partials, every
// workgroup reducing them the same way) and h = K_UU d_k; B takes alpha = <g, P g> / <d, h> and
// updates x, g and P g. d alternates between two buffers (A reads the previous one at stencil
// neighbours while it writes the new one). Workgroup 0 of each component records the scalars for
// the next kernel; a component that meets the stopping rule sets its done flag.
// ------------------------------------------------------------------------------------------------
struct CgState {
  double an;   // <g, P g> of the current iteration
  double spn;  // |g|^2
  double G0;   // |g0|^2 over D
  double gam, eta, del, xi;
  int it;
  int pad;
};

// Q1 stiffness row of node (x, y) over the undefined nodes only, u read through get(j).
template <typename Get>
__device__ __forceinline__ double stiff_row_u(const Get &get, const unsigned char *__restrict__ def, int x, int y,
                                              int n) {
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
        if (def[j]) continue;
        s += kl[bi][bj] * get(j);
      }
    }
  return s;
}

// Sum of cnt partials (every thread of the workgroup gets it; fixed order). Each thread loads a
// contiguous run of 8 before adding, so the loads of a sweep are in flight together.
__device__ __forceinline__ double wg_sum_partials(const double *__restrict__ p, int cnt, double *red) {
  double v = 0;
  for (int i0 = threadIdx.x * 8; i0 < cnt; i0 += 256 * 8) {
    double w[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) w[r] = (i0 + r < cnt) ? p[i0 + r] : 0.;
#pragma unroll
    for (int r = 0; r < 8; ++r) v += w[r];
  }
  __shared__ double out;
  const double s = block_sum<256>(v, red);
  if (threadIdx.x == 0) out = s;
  __syncthreads();
  return out;
}

// Start: on U g = K_UU x - rhs, P g, d = 0 - P g and the partials of <g,g>, <g,Pg>; on D the
// partial of |x - bc|^2. Grid (cdiv(n*n, 256), 2 components).
__global__ void k_cg_start(const double *__restrict__ x, const double *__restrict__ rhs,
                           const unsigned char *__restrict__ def, const double *__restrict__ invd,
                           double *__restrict__ g, double *__restrict__ hb, double *__restrict__ d,
                           double *__restrict__ part, int n, int nblk) {
  __shared__ double red[4];
  const int nn = n * n;
  const int c = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double gg = 0, gh = 0, g0 = 0;
  if (j < nn) {
    const int k = c * nn + j;
    if (def[j]) {
      const double v = x[k] - rhs[k];
      g0 = v * v;
    } else {
      const double *xc = x + c * nn;
      const double gv = stiff_row_u([&](int m) { return xc[m]; }, def, j % n, j / n, n) - rhs[k];
      const double pv = invd[j] * gv;
      g[k] = gv;
      hb[k] = pv;
      d[k] = 0.0 - pv;
      gg = gv * gv;
      gh = gv * pv;
    }
  }
  const double a = block_sum<256>(gg, red);
  const double b = block_sum<256>(gh, red);
  const double e = block_sum<256>(g0, red);
  if (threadIdx.x == 0) {
    part[(3 * c + 0) * nblk + blockIdx.x] = a;
    part[(3 * c + 1) * nblk + blockIdx.x] = b;
    part[(3 * c + 2) * nblk + blockIdx.x] = e;
  }
}

// Scalars of the start, one workgroup per component.
__global__ void k_cg_init(const double *__restrict__ part, int nblk, CgState *__restrict__ sb, int *__restrict__ done) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const double gg = wg_sum_partials(part + (3 * c + 0) * nblk, nblk, red);
  const double gh = wg_sum_partials(part + (3 * c + 1) * nblk, nblk, red);
  const double G0 = wg_sum_partials(part + (3 * c + 2) * nblk, nblk, red);
  if (threadIdx.x == 0) {
    CgState s;
    s.G0 = G0;
    s.gam = 1.0;   // g = g0 on D
    s.eta = 1.0;   // P g = g0
    s.del = -1.0;  // d = -P g
    s.xi = 0.0;
    s.spn = gg + G0;
    s.an = gh + G0;
    s.it = 0;
    s.pad = 0;
    sb[c] = s;
    done[c] = ((s.spn < 1e-16) || isnan(s.spn)) ? 1 : 0;
  }
}

// Iteration k, part A. Grid (nblkU, 2).
__global__ void k_cg_a(int k, const int *__restrict__ U, int nu, const unsigned char *__restrict__ def,
                       const double *__restrict__ dprev, double *__restrict__ dcur, const double *__restrict__ hb,
                       double *__restrict__ h2, const double *__restrict__ partb, double *__restrict__ parta, int nblk,
                       const CgState *__restrict__ sb, CgState *__restrict__ sa, int *__restrict__ done, int n) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  if (done[c]) return;
  const int nn = n * n;
  CgState s = sb[c];
  double beta = 0;
  if (k > 0) {
    // end of iteration k-1: <g, P g>, |g|^2, the stopping rule, then beta and the new direction
    const double bn = wg_sum_partials(partb + (2 * c + 0) * nblk, nblk, red) + (s.gam * s.eta) * s.G0;
    const double spn = wg_sum_partials(partb + (2 * c + 1) * nblk, nblk, red) + (s.gam * s.gam) * s.G0;
    beta = bn / s.an;
    s.an = bn;
    s.spn = spn;
    s.del = s.del * beta - s.eta;
  }
  s.it = k;
  if ((s.spn < 1e-16) || (k >= 1000) || isnan(s.spn)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sa[c] = s;
      done[c] = 1;
    }
    return;
  }
  const double *dp = dprev + c * nn, *hp = hb + c * nn;
  auto dk = [&](int m) { return k > 0 ? dp[m] * beta - hp[m] : dp[m]; };
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  double dh = 0;
  if (u < nu) {
    const int j = U[u];
    const double dv = dk(j);
    if (k > 0) dcur[c * nn + j] = dv;
    const double hv = stiff_row_u(dk, def, j % n, j / n, n);
    h2[c * nn + j] = hv;
    dh = dv * hv;
  }
  const double v = block_sum<256>(dh, red);
  if (threadIdx.x == 0) parta[c * nblk + blockIdx.x] = v;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.eta = s.del;  // h = A d is d on D
    sa[c] = s;
  }
}

// Iteration k, part B. Grid (nblkU, 2).
__global__ void k_cg_b(const int *__restrict__ U, int nu, double *__restrict__ x, const double *__restrict__ dcur,
                       double *__restrict__ g, double *__restrict__ hb, const double *__restrict__ h2,
                       const double *__restrict__ invd, const double *__restrict__ parta, double *__restrict__ partb,
                       int nblk, const CgState *__restrict__ sa, CgState *__restrict__ sb, const int *__restrict__ done,
                       int n) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  if (done[c]) return;
  const int nn = n * n;
  CgState s = sa[c];
  const double ad = wg_sum_partials(parta + c * nblk, nblk, red) + (s.del * s.eta) * s.G0;
  const double alpha = s.an / ad;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  double gh = 0, gg = 0;
  if (u < nu) {
    const int k = c * nn + U[u];
    x[k] += alpha * dcur[k];
    const double gv = g[k] + alpha * h2[k];
    g[k] = gv;
    const double pv = invd[k - c * nn] * gv;
    hb[k] = pv;
    gh = gv * pv;
    gg = gv * gv;
  }
  const double a = block_sum<256>(gh, red);
  const double b = block_sum<256>(gg, red);
  if (threadIdx.x == 0) {
    partb[(2 * c + 0) * nblk + blockIdx.x] = a;
    partb[(2 * c + 1) * nblk + blockIdx.x] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.xi += alpha * s.del;
    s.gam += alpha * s.eta;
    s.eta = s.gam;  // P g on D
    sb[c] = s;
  }
}

// x on D = x0 + xi g0 with g0 = x0 - bc, then the result in the field type.
template <typename D>
__global__ void k_cg_finish(const double *__restrict__ x, const double *__restrict__ rhs,
                            const unsigned char *__restrict__ def, const CgState *__restrict__ sb, D *__restrict__ out,
                            int n) {
  const int nn = n * n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * nn) return;
  const int c = i / nn, j = i - c * nn;
  double v = x[i];
  if (def[j]) v = v + sb[c].xi * (v - rhs[i]);
  out[i] = static_cast<D>(v);
}

// stable list of the undefined nodes (for hipcub::DeviceSelect::If over node indices)
struct UndefinedNode {
  const unsigned char *def;
  __host__ __device__ bool operator()(int j) const { return def[j] == 0; }
};

template <typename D>
__global__ void k_store(const double *__restrict__ x, D *__restrict__ out, int cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = static_cast<D>(x[i]);
}

}  // namespace msg
