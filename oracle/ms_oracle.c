/*
 * ms_oracle.c — FP64 CPU restatement of match-series' registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ms_oracle.h). It restates the reference's algorithm operation by
 * operation, in the reference's loop and summation order, so that it can serve as the checker of
 * the GPU kernels and as the timed single-thread CPU baseline. References are to QuocMesh
 * (quocmesh/modules/{aol,quoc}/ and projects/electronMicroscopy/).
 *
 * Build: make -C oracle lib   (gcc -std=c99 -O2 -ffp-contract=off)
 */
#include "ms_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(x, y, N) ((x) + (y) * (N))

/* 2-point Gauss rule on [0,1], tensorised; quad point k = (p[k/2], p[k%2]), weight 1/4
 * (baseFunctionSet.h:1535-1554 for the points, 1436-1467 for the ordering and weights). */
static const double QP[2] = {0.2113248654051871, 0.7886751345948129};
static const double QW = 0.25;

static double gridH(int N) { return 1.0 / (double)(N - 1); } /* rectangularGrid.h:50-57 */

/* Q1 basis on the unit square, local dofs (0,0),(1,0),(0,1),(1,1) (baseFunctionSet.h:127-193). */
static double basis(int b, double x, double y) {
  switch (b) {
    case 0: return (1.0 - x) * (1.0 - y);
    case 1: return x * (1.0 - y);
    case 2: return (1.0 - x) * y;
    default: return x * y;
  }
}
static double dxBasis(int b, double x, double y) {
  (void)x;
  switch (b) {
    case 0: return y - 1.0;
    case 1: return 1.0 - y;
    case 2: return -y;
    default: return y;
  }
}
static double dyBasis(int b, double x, double y) {
  (void)y;
  switch (b) {
    case 0: return x - 1.0;
    case 1: return -x;
    case 2: return 1.0 - x;
    default: return x;
  }
}
/* global index of local dof b of element (ex, ey) (configurators.h:880-890) */
static int l2g(int ex, int ey, int b, int N) { return (ex + ((b & 1) != 0)) + (ey + ((b & 2) != 0)) * N; }

/* cached basis values at the quad points (BaseFunctionSetInterface quad cache) */
static double BQ[4][4]; /* [b][q] */
static int bqInit = 0;
static void initBQ(void) {
  if (bqInit) return;
  for (int b = 0; b < 4; ++b)
    for (int q = 0; q < 4; ++q) BQ[b][q] = basis(b, QP[q / 2], QP[q % 2]);
  bqInit = 1;
}

/* ---------------------------------------------------------------------------------------------
 * Grid transfer
 * ------------------------------------------------------------------------------------------- */

/* CellCenteredRestrictOp::apply (cellCenteredGrid.h:165-180): 2x2 sum, x fastest, times 1/4. */
void msor_restrict(const double *fine, int N, double *coarse) {
  const int n = N / 2;
  const double scaling = 1.0 / 4.0;
  for (int cy = 0; cy < n; ++cy)
    for (int cx = 0; cx < n; ++cx) {
      double v = 0;
      v += fine[IDX(2 * cx, 2 * cy, N)];
      v += fine[IDX(2 * cx + 1, 2 * cy, N)];
      v += fine[IDX(2 * cx, 2 * cy + 1, N)];
      v += fine[IDX(2 * cx + 1, 2 * cy + 1, N)];
      coarse[IDX(cx, cy, n)] = scaling * v;
    }
}

/* CellCenteredProlongOp::apply (cellCenteredGrid.h:188-210): piecewise-constant injection. */
void msor_prolong(const double *coarse, int n, double *fine) {
  const int N = 2 * n;
  for (int cy = 0; cy < n; ++cy)
    for (int cx = 0; cx < n; ++cx) {
      const double v = coarse[IDX(cx, cy, n)];
      fine[IDX(2 * cx, 2 * cy, N)] = v;
      fine[IDX(2 * cx + 1, 2 * cy, N)] = v;
      fine[IDX(2 * cx, 2 * cy + 1, N)] = v;
      fine[IDX(2 * cx + 1, 2 * cy + 1, N)] = v;
    }
}

/* ScalarArray<QC_2D>::interpolate (scalarArray.h:499-529), including the Y == 0 branch. */
double msor_interpolate(const double *img, int N, double X, double Y) {
  int xL = (int)X, yL = (int)Y;
  if (X >= (double)(N - 1)) xL = N - 2;
  if (Y >= (double)(N - 1)) yL = N - 2;
  if (X < 0.) xL = 0;
  if (Y < 0.) yL = 0;
  X -= xL;
  Y -= yL;
  const int xU = xL + 1, yU = yL + 1;
  const double v0 = img[IDX(xL, yL, N)], v1 = img[IDX(xU, yL, N)];
  const double v2 = img[IDX(xL, yU, N)], v3 = img[IDX(xU, yU, N)];
  if (Y == 0) return (1 - X) * v0 + X * v1;
  return (1 - X) * (1 - Y) * v0 + X * (1 - Y) * v1 + (1 - X) * Y * v2 + X * Y * v3;
}

/* The same interpolation on a w x h array (ScalarArray<QC_2D>::interpolate with numX != numY) */
static double interpolateWH(const double *img, int w, int h, double X, double Y) {
  int xL = (int)X, yL = (int)Y;
  if (X >= (double)(w - 1)) xL = w - 2;
  if (Y >= (double)(h - 1)) yL = h - 2;
  if (X < 0.) xL = 0;
  if (Y < 0.) yL = 0;
  X -= xL;
  Y -= yL;
  const int xU = xL + 1, yU = yL + 1;
  const double v0 = img[xL + (size_t)yL * w], v1 = img[xU + (size_t)yL * w];
  const double v2 = img[xL + (size_t)yU * w], v3 = img[xU + (size_t)yU * w];
  if (Y == 0) return (1 - X) * v0 + X * v1;
  return (1 - X) * (1 - Y) * v0 + X * (1 - Y) * v1 + (1 - X) * Y * v2 + X * Y * v3;
}

void msor_resample_rect(const double *src, int w, int h, double *dst, int N) {
  const long double hx = (long double)((double)(w - 1) / (double)(N - 1));
  const long double hy = (long double)((double)(h - 1) / (double)(N - 1));
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) dst[IDX(x, y, N)] = interpolateWH(src, w, h, (double)(x * hx), (double)(y * hy));
}

/* ScalarArray<QC_2D>::resampleFrom via ResampleCopier (scalarArray.h:308-331,1052-1060): the step
 * (n-1)/(N-1) is held in long double; the scaled coordinate is rounded to double. */
void msor_resample(const double *src, int n, double *dst, int N) {
  /* ResampleCopier: the step divided in double, held in long double, products rounded to double */
  const long double hx = (long double)((double)(n - 1) / (double)(N - 1));
  const long double hy = hx;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      const double sx = (double)(x * hx), sy = (double)(y * hy);
      dst[IDX(x, y, N)] = msor_interpolate(src, n, sx, sy);
    }
}

static double rintd(double v) { return floor(v + 0.5); } /* aol::Rint */

/* DeformImage with ExtendWithConstant == true (deformations.h:62-101). */
void msor_deform_image(const double *img, const double *phi, int N, double ext, int nearest, double *out) {
  const double h = gridH(N);
  const int NN = N * N;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      const int i = IDX(x, y, N);
      double ds[2];
      int inside = 1;
      ds[0] = x + phi[i] / h;
      ds[1] = y + phi[NN + i] / h;
      for (int c = 0; c < 2; ++c) {
        if (ds[c] < 0.0 || ds[c] > (double)(N - 1) || isnan(ds[c])) inside = 0;
        if (nearest) ds[c] = rintd(ds[c]);
      }
      out[i] = inside ? msor_interpolate(img, N, ds[0], ds[1]) : ext;
    }
}

/* doFastDeformImage<QC_2D> (deformations.h:140-162): clamp to the domain, then interpolate. */
void msor_fast_deform_image(const double *img, const double *phi, int N, double *out) {
  const double h = gridH(N);
  const int NN = N * N;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      const int i = IDX(x, y, N);
      double dx = x + phi[i] / h, dy = y + phi[NN + i] / h;
      const double hi = (double)(N - 1);
      dx = dx < 0 ? 0 : (dx > hi ? hi : dx);
      dy = dy < 0 ? 0 : (dy > hi ? hi : dy);
      out[i] = msor_interpolate(img, N, dx, dy);
    }
}

/* ---------------------------------------------------------------------------------------------
 * SSD energy and its variation
 * ------------------------------------------------------------------------------------------- */

/* DiscreteFunctionDefault::evaluateAtQuadPoint (discreteFunction.h:239-246) */
static double evalAtQP(const double *dofs, int ex, int ey, int q, int N) {
  double w = 0.;
  for (int b = 0; b < 4; ++b) w += dofs[l2g(ex, ey, b, N)] * BQ[b][q];
  return w;
}

/* transformAndClipCoord with the per-axis flag (deformations.h:795-821) followed by
 * getLocalCoordsRegularRectangularGrid (configurators.h:32-50). */
static void transformAndClip(int ex, int ey, int q, double offx, double offy, int N, int *tel, double *tloc,
                             int *within) {
  const double h = gridH(N);
  double coord[2];
  coord[0] = ((double)ex + QP[q / 2]) * h;
  coord[1] = ((double)ey + QP[q % 2]) * h;
  coord[0] += offx;
  coord[1] += offy;
  for (int c = 0; c < 2; ++c) {
    const int width = (int)((N - 1) * h); /* = 1 for every dyadic N */
    within[c] = 1;
    if (coord[c] < 0.) {
      coord[c] = 0.;
      within[c] = 0;
    } else if (coord[c] >= width) {
      coord[c] = width - DBL_EPSILON;
      within[c] = 0;
    }
  }
  for (int c = 0; c < 2; ++c) {
    const double sc = coord[c] / h;
    int el = (short)sc;
    double loc = sc - el;
    if (el == N - 1 && fabs(loc - 0.0) < 1e-15) {
      el = N - 2;
      loc = 1;
    }
    tel[c] = el;
    tloc[c] = loc;
  }
}

/* DiscreteFunctionDefault::evaluate (discreteFunction.h:229-236) */
static double evalAt(const double *dofs, const int *el, const double *loc, int N) {
  double w = 0.;
  for (int b = 0; b < 4; ++b) w += dofs[l2g(el[0], el[1], b, N)] * basis(b, loc[0], loc[1]);
  return w;
}

/* DiscreteFunctionDefault::evaluateGradient (discreteFunction.h:259-268) */
static void evalGradAt(const double *dofs, const int *el, const double *loc, int N, double *g) {
  const double h = gridH(N);
  g[0] = 0;
  g[1] = 0;
  for (int b = 0; b < 4; ++b) {
    const double d = dofs[l2g(el[0], el[1], b, N)];
    double v0 = dxBasis(b, loc[0], loc[1]) / h, v1 = dyBasis(b, loc[0], loc[1]) / h;
    v0 *= d;
    v1 *= d;
    g[0] += v0;
    g[1] += v1;
  }
}

/* E_SSD = sum_El h^2 sum_q 1/4 * 1/2 (T(x_q + phi(x_q)) - R(x_q))^2
 * (SSDEnergy::evaluateIntegrand registration.h:1607-1615 in FENonlinIntegrationVectorInterface::applyAdd
 * FEOpInterface.h:2586-2613; elements x-fastest, rectangularGrid.h:139-150). */
double msor_ssd_energy(const double *R, const double *T, const double *phi, int N) {
  initBQ();
  const int NN = N * N;
  const double h = gridH(N), vol = h * h;
  double res = 0.;
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      double a = 0.;
      for (int q = 0; q < 4; ++q) {
        const double ox = evalAtQP(phi, ex, ey, q, N), oy = evalAtQP(phi + NN, ex, ey, q, N);
        int tel[2], within[2];
        double tloc[2];
        transformAndClip(ex, ey, q, ox, oy, N, tel, tloc, within);
        const double t = evalAt(T, tel, tloc, N) - evalAtQP(R, ex, ey, q, N);
        a += 0.5 * (t * t) * QW;
      }
      a *= vol;
      res += a;
    }
  return 0. + res;
}

/* SSDForce::getNonlinearity (registration.h:1646-1666) assembled by
 * FENonlinVectorOpInterface::applyAdd (FEOpInterface.h:1803-1853). */
void msor_ssd_force(const double *R, const double *T, const double *phi, int N, double *out) {
  initBQ();
  const int NN = N * N;
  const double h = gridH(N), vol = h * h;
  memset(out, 0, sizeof(double) * 2 * NN);
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      double nl[4][2];
      for (int q = 0; q < 4; ++q) {
        const double ox = evalAtQP(phi, ex, ey, q, N), oy = evalAtQP(phi + NN, ex, ey, q, N);
        int tel[2], within[2];
        double tloc[2];
        transformAndClip(ex, ey, q, ox, oy, N, tel, tloc, within);
        const double td = evalAt(T, tel, tloc, N);
        const double r = evalAtQP(R, ex, ey, q, N);
        double g[2];
        evalGradAt(T, tel, tloc, N, g);
        for (int c = 0; c < 2; ++c) nl[q][c] = within[c] ? g[c] * (td - r) : 0;
      }
      for (int dof = 0; dof < 4; ++dof) {
        double a0 = 0, a1 = 0;
        for (int q = 0; q < 4; ++q) {
          const double f = BQ[dof][q] * QW;
          a0 += f * nl[q][0];
          a1 += f * nl[q][1];
        }
        const int gi = l2g(ex, ey, dof, N);
        out[gi] += vol * a0;
        out[NN + gi] += vol * a1;
      }
    }
}

/* Assembled Q1 stiffness matrix applied to one component (StiffOp, FEOpInterface.h:2333-2348 with
 * the local matrix of FEOpInterface.h:1362-1398; MultilinFEBandMatrix<2D>). On the unit square
 * element the local matrix is (1/6)[4 -1 -1 -2; -1 4 -2 -1; -1 -2 4 -1; -2 -1 -1 4] in the dof
 * order (0,0),(1,0),(0,1),(1,1); the interior 9-point stencil is centre 8/3, neighbours -1/3.
 * Implemented row by row as the assembled band matrix: each row sums the element contributions. */
void msor_stiff_apply(const double *u, int N, double *out) {
  static const double KL[4][4] = {{4. / 6, -1. / 6, -1. / 6, -2. / 6},
                                  {-1. / 6, 4. / 6, -2. / 6, -1. / 6},
                                  {-1. / 6, -2. / 6, 4. / 6, -1. / 6},
                                  {-2. / 6, -1. / 6, -1. / 6, 4. / 6}};
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      /* assemble the row of node (x,y): coefficient of each of the 9 neighbours */
      double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int ey = y - 1; ey <= y; ++ey)
        for (int ex = x - 1; ex <= x; ++ex) {
          if (ex < 0 || ey < 0 || ex >= N - 1 || ey >= N - 1) continue;
          const int bi = (x - ex) + 2 * (y - ey);
          for (int bj = 0; bj < 4; ++bj) {
            const int jx = ex + (bj & 1), jy = ey + ((bj & 2) >> 1);
            c[jy - y + 1][jx - x + 1] += KL[bi][bj];
          }
        }
      double s = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int jx = x + dx, jy = y + dy;
          if (jx < 0 || jy < 0 || jx >= N || jy >= N) continue;
          s += c[dy + 1][dx + 1] * u[IDX(jx, jy, N)];
        }
      out[IDX(x, y, N)] = s;
    }
}

/* Assembled Q1 mass matrix applied to u (MassOp<ConfType>(grid, ASSEMBLED), FEOpInterface.h:2382-2397).
 * Local matrix of FELinScalarWeightedMassInterface::prepareLocalMatrix (FEOpInterface.h:1446-1475):
 * sum over the quadrature points of (b_i b_j * 1) * w_q, then times the element volume h^2. Rows are
 * assembled element by element and applied in column order, as for the stiffness matrix. */
void msor_mass_apply(const double *u, int N, double *out) {
  initBQ();
  const double h = gridH(N), vol = h * h;
  double ML[4][4] = {{0}};
  for (int q = 0; q < 4; ++q)
    for (int i = 0; i < 4; ++i) {
      const double bi = BQ[i][q];
      for (int j = i; j < 4; ++j) ML[i][j] += bi * BQ[j][q] * 1. * QW;
    }
  for (int i = 0; i < 4; ++i)
    for (int j = i; j < 4; ++j) ML[j][i] = ML[i][j] = ML[i][j] * vol;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int ey = y - 1; ey <= y; ++ey)
        for (int ex = x - 1; ex <= x; ++ex) {
          if (ex < 0 || ey < 0 || ex >= N - 1 || ey >= N - 1) continue;
          const int bi = (x - ex) + 2 * (y - ey);
          for (int bj = 0; bj < 4; ++bj) {
            const int jx = ex + (bj & 1), jy = ey + ((bj & 2) >> 1);
            c[jy - y + 1][jx - x + 1] += ML[bi][bj];
          }
        }
      double s = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int jx = x + dx, jy = y + dy;
          if (jx < 0 || jy < 0 || jx >= N || jy >= N) continue;
          s += c[dy + 1][dx + 1] * u[IDX(jx, jy, N)];
        }
      out[IDX(x, y, N)] = s;
    }
}

/* NormalizedCrossCorrelationEnergy::normalizeImageForNCC (registration.h:1934-1947): subtract the
 * mass-weighted mean, divide by sqrt(u.Mu) unless that is within 1e-15 of zero (appeqAbsolute,
 * aol.h:779-794). Returns sqrt(u.Mu). */
double msor_ncc_normalize(const double *img, int N, double *out) {
  const int NN = N * N;
  double *tmp = (double *)malloc(sizeof(double) * NN);
  msor_mass_apply(img, N, tmp);
  double mean = 0;
  for (int i = 0; i < NN; ++i) mean += tmp[i];                 /* Vector::sum, vec.cpp:847-870 */
  for (int i = 0; i < NN; ++i) out[i] = img[i];
  for (int i = 0; i < NN; ++i) out[i] += -mean;                /* addToAll, vec.h:633-639 */
  msor_mass_apply(out, N, tmp);
  double dot = 0;
  for (int i = 0; i < NN; ++i) dot += out[i] * tmp[i];         /* ScalarProduct, vec.cpp:33-40 */
  const double var = sqrt(dot);
  if (!(fabs(var - 0.) < 1e-15))
    for (int i = 0; i < NN; ++i) out[i] /= var;                /* vec.cpp:729-734 */
  free(tmp);
  return var;
}

/* NormalizedCrossCorrelationEnergy::computeEnergy (registration.h:1953-1964) with the reference
 * already normalised: -(M Tn) . Rn, Tn the normalised FastDeformImage of T. Writes Tn and returns
 * the energy; *var_t receives the variance of the deformed template. */
static double nccEnergyNormalized(const double *Rn, const double *T, const double *phi, int N, double *Tn,
                                  double *var_t) {
  const int NN = N * N;
  double *def = (double *)malloc(sizeof(double) * NN);
  msor_fast_deform_image(T, phi, N, def);
  *var_t = msor_ncc_normalize(def, N, Tn);
  msor_mass_apply(Tn, N, def);
  double dot = 0;
  for (int i = 0; i < NN; ++i) dot += def[i] * Rn[i];
  free(def);
  return -dot;
}

double msor_ncc_energy(const double *R, const double *T, const double *phi, int N) {
  const int NN = N * N;
  double *Rn = (double *)malloc(sizeof(double) * NN), *Tn = (double *)malloc(sizeof(double) * NN);
  double var_t;
  msor_ncc_normalize(R, N, Rn);
  const double e = nccEnergyNormalized(Rn, T, phi, N, Tn, &var_t);
  free(Rn);
  free(Tn);
  return e;
}

/* transformCoord<ConfType, false> (deformations.h:623-668, 681-693): grid coordinates of the quad
 * point plus offset/h; false when the point leaves the grid on the first axis tested. */
static int transformNoClip(int ex, int ey, int q, double offx, double offy, int N, int *tel, double *tloc) {
  const double h = gridH(N);
  double coord[2];
  coord[0] = (double)ex + QP[q / 2];
  coord[1] = (double)ey + QP[q % 2];
  const double off[2] = {offx, offy};
  for (int i = 0; i < 2; ++i) {
    const double tc = coord[i] + off[i] / h;
    int el = (short)tc;
    double loc = tc - el;
    if (el == N - 1 && loc == 0) {
      el = N - 2;
      loc = 1;
    }
    tel[i] = el;
    tloc[i] = loc;
    if (tc < 0. || el < 0 || el >= N - 1) return 0;
  }
  return 1;
}

/* NormalizedCrossCorrelationEnergy::applyDerivative (registration.h:1974-1984): w = (Rn + E Tn)/var_T
 * (no division when var_T is within 1e-15 of zero), then CrossCorrelationForce(w, T)
 * (registration.h:1826-1862): -w(x_q) grad T(x_q + phi(x_q)) at every quad point that stays in
 * the grid, zero elsewhere, assembled by FENonlinVectorOpInterface::applyAdd (FEOpInterface.h:1803-1853). */
static void nccForceNormalized(const double *Rn, const double *T, const double *phi, int N, double *out) {
  initBQ();
  const int NN = N * N;
  const double h = gridH(N), vol = h * h;
  double *Tn = (double *)malloc(sizeof(double) * NN), *w = (double *)malloc(sizeof(double) * NN);
  double var_t;
  const double e = nccEnergyNormalized(Rn, T, phi, N, Tn, &var_t);
  for (int i = 0; i < NN; ++i) w[i] = Rn[i];
  for (int i = 0; i < NN; ++i) w[i] += e * Tn[i];              /* addMultiple, vec.h */
  if (!(fabs(var_t - 0.) < 1e-15))
    for (int i = 0; i < NN; ++i) w[i] /= var_t;
  memset(out, 0, sizeof(double) * 2 * NN);
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      double nl[4][2];
      for (int q = 0; q < 4; ++q) {
        const double ox = evalAtQP(phi, ex, ey, q, N), oy = evalAtQP(phi + NN, ex, ey, q, N);
        int tel[2];
        double tloc[2];
        if (!transformNoClip(ex, ey, q, ox, oy, N, tel, tloc)) {
          nl[q][0] = nl[q][1] = 0;
          continue;
        }
        double g[2];
        evalGradAt(T, tel, tloc, N, g);
        const double wq = -evalAtQP(w, ex, ey, q, N);
        nl[q][0] = g[0] * wq;
        nl[q][1] = g[1] * wq;
      }
      for (int dof = 0; dof < 4; ++dof) {
        double a0 = 0, a1 = 0;
        for (int q = 0; q < 4; ++q) {
          const double f = BQ[dof][q] * QW;
          a0 += f * nl[q][0];
          a1 += f * nl[q][1];
        }
        const int gi = l2g(ex, ey, dof, N);
        out[gi] += vol * a0;
        out[NN + gi] += vol * a1;
      }
    }
  free(Tn);
  free(w);
}

void msor_ncc_force(const double *R, const double *T, const double *phi, int N, double *out) {
  double *Rn = (double *)malloc(sizeof(double) * N * N);
  msor_ncc_normalize(R, N, Rn);
  nccForceNormalized(Rn, T, phi, N, out);
  free(Rn);
}

/* DisplacementLengthEnergy::apply (hyperelastic.h:178-183): 1/2 sum_c phi_c . K phi_c, the
 * MultiVector dot product summed component by component (multiVector.h:570-578, vec.cpp:33-40). */
double msor_reg_energy(const double *phi, int N) {
  const int NN = N * N;
  double *k = (double *)malloc(sizeof(double) * NN);
  double dot = 0;
  for (int c = 0; c < 2; ++c) {
    msor_stiff_apply(phi + c * NN, N, k);
    double s = 0;
    for (int i = 0; i < NN; ++i) s += phi[c * NN + i] * k[i];
    dot += s;
  }
  free(k);
  return 0.5 * dot;
}

/* ---------------------------------------------------------------------------------------------
 * H1 metric: FD heat-equation multigrid (linearSmoothOp.cpp)
 * ------------------------------------------------------------------------------------------- */

/* HeatEquation2DFD::apply (linearSmoothOp.cpp:59-146): u + tau (N-1)^2 (4u - sum of neighbours),
 * with mirrored (doubled) neighbours on edges and corners. */
void msor_heat_apply(const double *u, int N, double tau, double *out) {
  if (N == 1) {
    out[0] = u[0];
    return;
  }
  const double hs = tau * (double)((N - 1) * (N - 1));
  const int off[4] = {1, -1, N, -N};
  for (int X = 1; X < N - 1; ++X)
    for (int Y = 1; Y < N - 1; ++Y) {
      const int i = IDX(X, Y, N);
      const double a = u[i] * 4;
      double b = 0.0;
      for (int k = 0; k < 4; ++k) b -= u[i + off[k]];
      out[i] = u[i] + hs * (a + b);
    }
  const int e = N - 1;
  const int corners[4][2] = {{0, 0}, {e, 0}, {0, e}, {e, e}};
  const int coffs[4][2] = {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  for (int k = 0; k < 4; ++k) {
    const int cx = corners[k][0], cy = corners[k][1];
    const double a = u[IDX(cx, cy, N)] * 4;
    const double b = -u[IDX(cx + coffs[k][0], cy, N)] - u[IDX(cx, cy + coffs[k][1], N)];
    out[IDX(cx, cy, N)] = u[IDX(cx, cy, N)] + hs * (a + 2 * b);
  }
  for (int X = 1; X < N - 1; ++X) {
    double a = u[IDX(X, e, N)] * 4;
    double b = -2 * (u[IDX(X, e - 1, N)]);
    b -= u[IDX(X + 1, e, N)];
    b -= u[IDX(X - 1, e, N)];
    out[IDX(X, e, N)] = u[IDX(X, e, N)] + hs * (a + b);
    a = u[IDX(X, 0, N)] * 4;
    b = -2 * (u[IDX(X, 1, N)]);
    b -= u[IDX(X + 1, 0, N)];
    b -= u[IDX(X - 1, 0, N)];
    out[IDX(X, 0, N)] = u[IDX(X, 0, N)] + hs * (a + b);
  }
  for (int X = 1; X < N - 1; ++X) {
    double a = u[IDX(e, X, N)] * 4;
    double b = -2 * (u[IDX(e - 1, X, N)]);
    b -= u[IDX(e, X + 1, N)];
    b -= u[IDX(e, X - 1, N)];
    out[IDX(e, X, N)] = u[IDX(e, X, N)] + hs * (a + b);
    a = u[IDX(0, X, N)] * 4;
    b = -2 * (u[IDX(1, X, N)]);
    b -= u[IDX(0, X + 1, N)];
    b -= u[IDX(0, X - 1, N)];
    out[IDX(0, X, N)] = u[IDX(0, X, N)] + hs * (a + b);
  }
}

/* HeatMultigrid::jacobi (linearSmoothOp.cpp:556-580): x += (rhs - A x) / (1 + 4 tau (N-1)^2). */
static void jacobi(int N, double tau, double *x, const double *rhs, double *dummy) {
  const double diag = 1 + 4 * tau * (double)((N - 1) * (N - 1));
  const double relax = 1.0;
  msor_heat_apply(x, N, tau, dummy);
  for (int l = 0; l < N * N; ++l) {
    const double v = rhs[l] - dummy[l];
    x[l] += relax * v / diag;
  }
}

/* MDMultigrid::solve (linearSmoothOp.cpp:451-494): 2 pre-smoothing Jacobi sweeps, residual,
 * restriction, recursion, injection of the coarse correction, 2 post-smoothing sweeps;
 * 100 sweeps on the 1x1 level. */
static void mgSolve(int level, double tau, double **x, double **rhs, double **dummy, double **res) {
  const int N = 1 << level;
  if (level == 0) {
    for (int i = 0; i < 100; ++i) jacobi(N, tau, x[0], rhs[0], dummy[0]);
    return;
  }
  for (int i = 0; i < 2; ++i) jacobi(N, tau, x[level], rhs[level], dummy[level]);
  msor_heat_apply(x[level], N, tau, dummy[level]);
  for (int l = 0; l < N * N; ++l) res[level][l] = rhs[level][l] - dummy[level][l];
  msor_restrict(res[level], N, rhs[level - 1]);
  memset(x[level - 1], 0, sizeof(double) * (N / 2) * (N / 2));
  mgSolve(level - 1, tau, x, rhs, dummy, res);
  msor_prolong(x[level - 1], N / 2, dummy[level]);
  for (int l = 0; l < N * N; ++l) x[level][l] += dummy[level][l];
  for (int i = 0; i < 2; ++i) jacobi(N, tau, x[level], rhs[level], dummy[level]);
}

/* LinearSmoothOp::apply (linearSmoothOp.cpp:591-612) with tau = sigma^2/2 (linearSmoothOp.h:50-52). */
void msor_vcycle(const double *d, int N, double sigma, double *out) {
  int L = 0;
  while ((1 << L) < N) ++L;
  const double tau = 0.5 * (sigma * sigma);
  double *x[32], *rhs[32], *dummy[32], *res[32];
  for (int l = 0; l <= L; ++l) {
    const size_t n = (size_t)1 << (2 * l);
    x[l] = (double *)calloc(n, sizeof(double));
    rhs[l] = (double *)calloc(n, sizeof(double));
    dummy[l] = (double *)calloc(n, sizeof(double));
    res[l] = (double *)calloc(n, sizeof(double));
  }
  memcpy(rhs[L], d, sizeof(double) * N * N);
  mgSolve(L, tau, x, rhs, dummy, res);
  memcpy(out, x[L], sizeof(double) * N * N);
  for (int l = 0; l <= L; ++l) {
    free(x[l]);
    free(rhs[l]);
    free(dummy[l]);
    free(res[l]);
  }
}

/* ---------------------------------------------------------------------------------------------
 * Consistency energy of reduceDeformations (AmbrosioTortorelli.h:680-800)
 * ------------------------------------------------------------------------------------------- */

/* ConsistencyEnergyOp::evaluateIntegrand (AmbrosioTortorelli.h:705-728) integrated as
 * FENonlinIntegrationVectorInterface::applyAdd: 1/2 int |psi(x + phi(x)) + phi(x)|^2 with the clipped
 * deformed quad point of transformAndClipCoord (deformations.h:764-787). */
double msor_consistency_energy(const double *psi, const double *phi, int N) {
  initBQ();
  const int NN = N * N;
  const double h = gridH(N), vol = h * h;
  double res = 0.;
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      double a = 0.;
      for (int q = 0; q < 4; ++q) {
        const double off[2] = {evalAtQP(phi, ex, ey, q, N), evalAtQP(phi + NN, ex, ey, q, N)};
        int tel[2], within[2];
        double tloc[2];
        transformAndClip(ex, ey, q, off[0], off[1], N, tel, tloc, within);
        double energy = 0.;
        for (int i = 0; i < 2; ++i) {
          const double v = evalAt(psi + i * NN, tel, tloc, N) + off[i];
          energy += v * v;
        }
        a += 0.5 * energy * QW;
      }
      a *= vol;
      res += a;
    }
  return 0. + res;
}

/* VariationOfConsistencyEnergyOp::getNonlinearity (AmbrosioTortorelli.h:755-799): (Dpsi + I)^T
 * (psi + offset) at the clipped deformed quad point, zero per clipped axis; assembled as
 * FENonlinVectorOpInterface::applyAdd. */
void msor_consistency_force(const double *psi, const double *phi, int N, double *out) {
  initBQ();
  const int NN = N * N;
  const double h = gridH(N), vol = h * h;
  memset(out, 0, sizeof(double) * 2 * NN);
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      double nl[4][2];
      for (int q = 0; q < 4; ++q) {
        const double off[2] = {evalAtQP(phi, ex, ey, q, N), evalAtQP(phi + NN, ex, ey, q, N)};
        int tel[2], within[2];
        double tloc[2];
        transformAndClip(ex, ey, q, off[0], off[1], N, tel, tloc, within);
        double v[2], J[2][2];
        for (int c = 0; c < 2; ++c) {
          v[c] = evalAt(psi + c * NN, tel, tloc, N);
          evalGradAt(psi + c * NN, tel, tloc, N, J[c]);
        }
        for (int c = 0; c < 2; ++c) v[c] += off[c];
        for (int c = 0; c < 2; ++c) J[c][c] += 1.;
        for (int i = 0; i < 2; ++i) {                   /* Mat::multTransposed, smallMat.h:358-364 */
          nl[q][i] = 0;
          for (int j = 0; j < 2; ++j) nl[q][i] += J[j][i] * v[j];
        }
        for (int i = 0; i < 2; ++i)
          if (!within[i]) nl[q][i] = 0;
      }
      for (int dof = 0; dof < 4; ++dof) {
        double a0 = 0, a1 = 0;
        for (int q = 0; q < 4; ++q) {
          const double f = BQ[dof][q] * QW;
          a0 += f * nl[q][0];
          a1 += f * nl[q][1];
        }
        const int gi = l2g(ex, ey, dof, N);
        out[gi] += vol * a0;
        out[NN + gi] += vol * a1;
      }
    }
}

/* ---------------------------------------------------------------------------------------------
 * Seismic x-shift registration (matchSeismicSeries): a displacement along x only, phi_x, whose data
 * term is the mean over the image rows ("lines") of the 1-D normalised cross-correlation of the
 * template row shifted by the row of phi_x against the reference row (FELineShiftRegisEnergy,
 * registration.h:2137-2193, with NCCRegistrationConfigurator on the 1-D grid of the row), and whose
 * regulariser is lambda/2 phi_x^T (K_xx + yRegFactor K_yy) phi_x (findXShift, matchSeries.h:805-868).
 * The 1-D configurator integrates with 2 Gauss points of weight 1/2 (baseFunctionSet.h:1535-1554).
 * ------------------------------------------------------------------------------------------- */

/* aol::Vector::interpolate (vec.h:461-471) */
static double lineInterp(const double *v, int n, double pos) {
  const int p = (int)pos;
  const double pf = pos - p;
  if (p == n - 1) return v[p];
  return (1 - pf) * v[p] + pf * v[p + 1];
}

/* The assembled 1-D mass matrix (MassOp ASSEMBLED): element local entries by the 2-point rule, added
 * in element order; diag[i] = M_ii, off[i] = M_i,i+1 = M_i+1,i. */
static void mass1D(int n, double *diag, double *off) {
  const double h = gridH(n);
  double L[2][2] = {{0, 0}, {0, 0}};
  for (int q = 0; q < 2; ++q) {
    const double b[2] = {1.0 - QP[q], QP[q]};
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) L[a][c] += 0.5 * b[a] * b[c];
  }
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) L[a][c] *= h;
  for (int i = 0; i < n; ++i) diag[i] = 0;
  for (int e = 0; e < n - 1; ++e) {
    diag[e] += L[0][0];
    off[e] = L[0][1];
    diag[e + 1] += L[1][1];
  }
}

/* y = M x, each row summed in ascending column order */
static void massApply1D(const double *x, int n, const double *diag, const double *off, double *y) {
  for (int i = 0; i < n; ++i) {
    double s = 0;
    if (i > 0) s += off[i - 1] * x[i - 1];
    s += diag[i] * x[i];
    if (i < n - 1) s += off[i] * x[i + 1];
    y[i] = s;
  }
}

/* normalizeImageForNCC (registration.h:1934-1947) on one row; returns sqrt(u.Mu) */
static double nccNormalize1D(const double *img, int n, const double *diag, const double *off, double *out,
                             double *tmp) {
  massApply1D(img, n, diag, off, tmp);
  double mean = 0;
  for (int i = 0; i < n; ++i) mean += tmp[i];
  for (int i = 0; i < n; ++i) out[i] = img[i];
  for (int i = 0; i < n; ++i) out[i] += -mean;
  massApply1D(out, n, diag, off, tmp);
  double dot = 0;
  for (int i = 0; i < n; ++i) dot += out[i] * tmp[i];
  const double var = sqrt(dot);
  if (!(fabs(var - 0.) < 1e-15))
    for (int i = 0; i < n; ++i) out[i] /= var;
  return var;
}

/* One row: FastDeformImage<1-D> (deformations.h:111-137: t at i + y_i/h clamped to [0, n-1]),
 * normalisation, E = -(M tn) . rn. Writes tn and the deformed row's deviation. */
static double lineNcc(const double *rn, const double *t, const double *y, int n, const double *diag,
                      const double *off, double *tn, double *var_t, double *tmp, double *def) {
  const double h = gridH(n);
  for (int i = 0; i < n; ++i) {
    double ds = i + y[i] / h;
    ds = ds < 0 ? 0 : (ds > (double)(n - 1) ? (double)(n - 1) : ds);
    def[i] = lineInterp(t, n, ds);
  }
  *var_t = nccNormalize1D(def, n, diag, off, tn, tmp);
  massApply1D(tn, n, diag, off, def);
  double dot = 0;
  for (int i = 0; i < n; ++i) dot += def[i] * rn[i];
  return -dot;
}

/* The assembled K_xx + yRegFactor K_yy of findXShift (FEOpMixedDerivative(0,0) and (1,1) added into one
 * matrix, the second scaled, matchSeries.h:820-826) applied to u: element local matrices by the 2x2
 * rule, added per element in element order, each row summed over its 3x3 neighbours. */
static void xshiftStiffApply(const double *u, int N, double yreg, double *out) {
  initBQ();
  double Kx[4][4] = {{0}}, Ky[4][4] = {{0}};
  for (int q = 0; q < 4; ++q) {
    const double qx = QP[q / 2], qy = QP[q % 2];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        Kx[a][b] += dxBasis(a, qx, qy) * dxBasis(b, qx, qy) * QW;
        Ky[a][b] += dyBasis(a, qx, qy) * dyBasis(b, qx, qy) * QW;
      }
  }
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int ey = y - 1; ey <= y; ++ey)
        for (int ex = x - 1; ex <= x; ++ex) {
          if (ex < 0 || ey < 0 || ex >= N - 1 || ey >= N - 1) continue;
          const int bi = (x - ex) + 2 * (y - ey);
          for (int bj = 0; bj < 4; ++bj) {
            const int jx = ex + (bj & 1), jy = ey + ((bj & 2) >> 1);
            c[jy - y + 1][jx - x + 1] += Kx[bi][bj];
          }
        }
      for (int ey = y - 1; ey <= y; ++ey)
        for (int ex = x - 1; ex <= x; ++ex) {
          if (ex < 0 || ey < 0 || ex >= N - 1 || ey >= N - 1) continue;
          const int bi = (x - ex) + 2 * (y - ey);
          for (int bj = 0; bj < 4; ++bj) {
            const int jx = ex + (bj & 1), jy = ey + ((bj & 2) >> 1);
            c[jy - y + 1][jx - x + 1] += yreg * Ky[bi][bj];
          }
        }
      double s = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int jx = x + dx, jy = y + dy;
          if (jx < 0 || jy < 0 || jx >= N || jy >= N) continue;
          s += c[dy + 1][dx + 1] * u[IDX(jx, jy, N)];
        }
      out[IDX(x, y, N)] = s;
    }
}

/* E_data = (sum over rows whose reference and template both start with a number of E_row) / N, rows
 * summed in order (FELineShiftRegisEnergy::apply); E_reg = 1/2 phi_x . A phi_x (QuadraticFormOp,
 * op.h:1417-1421). */
void msor_xshift_energy(const double *R, const double *T, const double *phix, int N, double yreg, double *e_data,
                        double *e_reg) {
  double *diag = (double *)malloc(sizeof(double) * N), *off = (double *)malloc(sizeof(double) * N);
  double *rn = (double *)malloc(sizeof(double) * N), *tn = (double *)malloc(sizeof(double) * N);
  double *tmp = (double *)malloc(sizeof(double) * N), *def = (double *)malloc(sizeof(double) * N);
  double *au = (double *)malloc(sizeof(double) * N * N);
  mass1D(N, diag, off);
  double E = 0;
  for (int line = 0; line < N; ++line) {
    const double *r = R + (size_t)line * N, *t = T + (size_t)line * N;
    if (isnan(r[0]) || isnan(t[0])) continue;
    nccNormalize1D(r, N, diag, off, rn, tmp);
    double var_t;
    E += lineNcc(rn, t, phix + (size_t)line * N, N, diag, off, tn, &var_t, tmp, def);
  }
  E /= N;
  if (e_data) *e_data = E;
  if (e_reg) {
    xshiftStiffApply(phix, N, yreg, au);
    double dot = 0;
    for (int i = 0; i < N * N; ++i) dot += au[i] * phix[i];
    *e_reg = 0.5 * dot;
  }
  free(diag); free(off); free(rn); free(tn); free(tmp); free(def); free(au);
}

/* DE_data: per row the derivative of its NCC energy (VariationOfNormalizedCrossCorrelationEnergy:
 * w = (rn + E tn) / var_t, then CrossCorrelationForce on the row: -w(x_q) T'(x_q + y(x_q)) at each
 * quad point whose shifted position stays on the row, zero elsewhere), divided by N; plus
 * lambda A phi_x. out: N*N values (the x component). */
void msor_xshift_gradient(const double *R, const double *T, const double *phix, int N, double yreg, double lambda,
                          double *out) {
  const double h = gridH(N);
  double *diag = (double *)malloc(sizeof(double) * N), *off = (double *)malloc(sizeof(double) * N);
  double *rn = (double *)malloc(sizeof(double) * N), *tn = (double *)malloc(sizeof(double) * N);
  double *tmp = (double *)malloc(sizeof(double) * N), *def = (double *)malloc(sizeof(double) * N);
  double *w = (double *)malloc(sizeof(double) * N), *au = (double *)malloc(sizeof(double) * N * N);
  mass1D(N, diag, off);
  memset(out, 0, sizeof(double) * N * N);
  for (int line = 0; line < N; ++line) {
    const double *r = R + (size_t)line * N, *t = T + (size_t)line * N, *y = phix + (size_t)line * N;
    if (isnan(r[0]) || isnan(t[0])) continue;
    nccNormalize1D(r, N, diag, off, rn, tmp);
    double var_t;
    const double e = lineNcc(rn, t, y, N, diag, off, tn, &var_t, tmp, def);
    for (int i = 0; i < N; ++i) w[i] = rn[i];
    for (int i = 0; i < N; ++i) w[i] += e * tn[i];
    if (!(fabs(var_t - 0.) < 1e-15))
      for (int i = 0; i < N; ++i) w[i] /= var_t;
    double *o = out + (size_t)line * N;
    for (int el = 0; el < N - 1; ++el) {
      double nl[2];
      for (int q = 0; q < 2; ++q) {
        const double xi = QP[q];
        const double off_q = (1 - xi) * y[el] + xi * y[el + 1];       /* evaluateAtQuadPoint */
        const double tc = (double)el + xi + off_q / h;                  /* transformCoord<.., false> */
        int tel = (short)tc;
        double tloc = tc - tel;
        if (tel == N - 1 && tloc == 0) {
          tel = N - 2;
          tloc = 1;
        }
        if (tc < 0. || tel < 0 || tel >= N - 1) {
          nl[q] = 0;
          continue;
        }
        const double g = (t[tel + 1] - t[tel]) / h;                     /* evaluateGradient */
        const double wq = (1 - xi) * w[el] + xi * w[el + 1];
        nl[q] = g * -wq;
      }
      for (int dof = 0; dof < 2; ++dof) {
        double a = 0;
        for (int q = 0; q < 2; ++q) a += (dof ? QP[q] : 1 - QP[q]) * 0.5 * nl[q];
        o[el + dof] += h * a;
      }
    }
    for (int i = 0; i < N; ++i) o[i] /= N;
  }
  xshiftStiffApply(phix, N, yreg, au);
  for (int i = 0; i < N * N; ++i) out[i] += lambda * au[i];
  free(diag); free(off); free(rn); free(tn); free(tmp); free(def); free(w); free(au);
}

/* ---------------------------------------------------------------------------------------------
 * One level of the H1 gradient flow with nonlinear CG and Armijo step control
 * ------------------------------------------------------------------------------------------- */

typedef struct {
  const double *R, *T;
  int N;
  double lambda;
  double *tmp; /* 2*N*N scratch for positions */
  int evals;
  int data_term;
  double *Rn, *Tn; /* NCC: the normalised reference (set once per level, as the reference's
                      energy object does in its constructor) and a scratch image */
  const double *psi; /* consistency: the n deformations, 2*N*N each */
  int npsi;
  double yreg; /* x-shift: yRegFactor */
} Energy;

enum { DT_CONSISTENCY = 100 }; /* internal: reduceDeformations' energy */

/* E = (0 + E_data) + lambda * E_reg (LinCombOp::applyAdd, op.h:255-272) */
static double energyAt(Energy *e, const double *phi) {
  double E = 0;
  if (e->data_term == DT_CONSISTENCY) {  /* LinCombOp of the n consistency energies, coefficient 1 */
    for (int i = 0; i < e->npsi; ++i) E += msor_consistency_energy(e->psi + (size_t)i * 2 * e->N * e->N, phi, e->N);
    return E;
  }
  if (e->data_term == MSOR_XSHIFT) {  /* (0 + E_data) + lambda E_reg (LinCombOp, matchSeries.h:844-846) */
    double ed, er;
    msor_xshift_energy(e->R, e->T, phi, e->N, e->yreg, &ed, &er);
    E += ed;
    E += e->lambda * er;
    return E;
  }
  if (e->data_term == MSOR_NCC) {
    double var_t;
    E += nccEnergyNormalized(e->Rn, e->T, phi, e->N, e->Tn, &var_t);
  } else {
    E += msor_ssd_energy(e->R, e->T, phi, e->N);
  }
  E += e->lambda * msor_reg_energy(phi, e->N);
  return E;
}

/* DE = (0 + DE_data) + lambda * K phi, componentwise */
static void gradientAt(Energy *e, const double *phi, double *out) {
  const int NN = e->N * e->N;
  if (e->data_term == DT_CONSISTENCY) {  /* Dest = 0; Dest += tmp_i (LinCombOp::applyAdd, op.h:254-272) */
    double *tmp = (double *)malloc(sizeof(double) * 2 * NN);
    memset(out, 0, sizeof(double) * 2 * NN);
    for (int i = 0; i < e->npsi; ++i) {
      msor_consistency_force(e->psi + (size_t)i * 2 * NN, phi, e->N, tmp);
      for (int k = 0; k < 2 * NN; ++k) out[k] += tmp[k];
    }
    free(tmp);
    return;
  }
  if (e->data_term == MSOR_XSHIFT) {  /* the y component stays zero */
    msor_xshift_gradient(e->R, e->T, phi, e->N, e->yreg, e->lambda, out);
    memset(out + NN, 0, sizeof(double) * NN);
    return;
  }
  double *k = (double *)malloc(sizeof(double) * NN);
  if (e->data_term == MSOR_NCC)
    nccForceNormalized(e->Rn, e->T, phi, e->N, out);
  else
    msor_ssd_force(e->R, e->T, phi, e->N, out);
  for (int c = 0; c < 2; ++c) {
    msor_stiff_apply(phi + c * NN, e->N, k);
    for (int i = 0; i < NN; ++i) out[c * NN + i] += e->lambda * k[i];
  }
  free(k);
}

/* ArmijoLineSearchHelpFunction_evaluate (gradientDescent.h:154-174): E(x + tau d) */
static double evalStep(Energy *e, const double *x, const double *d, double tau) {
  const int n = 2 * e->N * e->N;
  for (int i = 0; i < n; ++i) e->tmp[i] = x[i] + tau * d[i];
  e->evals++;
  return energyAt(e, e->tmp);
}

static double dot2(const double *a, const double *b, int NN) {
  double dot = 0;
  for (int c = 0; c < 2; ++c) {
    double s = 0;
    for (int i = 0; i < NN; ++i) s += a[c * NN + i] * b[c * NN + i];
    dot += s;
  }
  return dot;
}

static const double TAU_MIN = 9.313225746154785e-10; /* 2^-30 (ArmijoSearch.h:64-73) */
static const double TAU_MAX = 33554432.0;            /* 2^25 */
static const double ARMIJO_SIGMA = 0.5;

/* getTimestepWidthWithArmijoLineSearch (ArmijoSearch.h:118-200) */
static double armijo(Energy *e, const double *x, const double *d, const double *rawDE, double oldTau, double f) {
  const int NN = e->N * e->N;
  double tau = oldTau < TAU_MIN ? TAU_MIN : (oldTau > TAU_MAX ? TAU_MAX : oldTau);
  double fNew = evalStep(e, x, d, tau);
  const double Df = dot2(rawDE, d, NN);
  double G = (fNew - f) / (Df * tau);
  if (Df >= 0) return 0;
  if (!isfinite(f)) return 0;
  if (G >= ARMIJO_SIGMA && f >= fNew) {
    do {
      tau *= 2.;
      fNew = evalStep(e, x, d, tau);
      G = (fNew - f) / (Df * tau);
    } while (G >= ARMIJO_SIGMA && f >= fNew && tau <= TAU_MAX);
    tau *= 0.5;
  } else {
    do {
      if (tau > TAU_MIN) tau *= 0.5;
      fNew = evalStep(e, x, d, tau);
      G = (fNew - f) / (Df * tau);
    } while ((!isfinite(fNew) || (G < ARMIJO_SIGMA || f < fNew)) && (tau > TAU_MIN));
  }
  return tau > TAU_MIN ? tau : 0.;
}

/* GradientDescentBase::applySingle (gradientDescent.h:295-464) as configured by
 * StandardRegistration::updateTransformationWithGD (registration.h:1003-1021): H1 metric
 * (H1GradientDescent::smoothDirection, gradientDescent.h:652-662), Fletcher-Reeves CG restarted every
 * 10 steps (calcTaudAndDescentDir, L200-236), automatic filter width starting at 1 and halved when
 * the energy decrease drops below 1e-5 (GradientDescentWithAutomaticFilterWidth L675-691; L416-439).
 * An energy output stream is attached in the reference run, so the energy at the new position is
 * re-evaluated after each update instead of being taken from the Armijo cache (L361-396). */
/* The descent loop of GradientDescentBase::applySingle for an energy e, from phi (updated in place).
 * auto_sigma: GradientDescentWithAutomaticFilterWidth's schedule from sigma 1, else sigma0 fixed
 * (H1GradientDescent alone, gradientDescent.h:639-650). grad_stop: USE_GRADIENT_BASED_STOPPING
 * (continue while |DE at the old position| > stop_eps, L416), else the energy decrease. */
static int descent(Energy *e, double *phi, int max_iter, double stop_eps, double sigma0, int auto_sigma,
                   int grad_stop, double *energy_trace, double *tau_trace, double *sigma_trace, int *evals_trace,
                   int trace_cap, int *n_armijo_calls, double *final_energy) {
  const int N = e->N, NN = N * N, n2 = 2 * NN;
  double *de = (double *)malloc(sizeof(double) * n2);
  double *dir = (double *)malloc(sizeof(double) * n2);
  double *oldDir = (double *)calloc(n2, sizeof(double));
  double *oldGrad = (double *)calloc(n2, sizeof(double));
  double tau = 1.0;
  double filterWidth = auto_sigma ? 1.0 : sigma0;
  const double stopFilterWidth = 0.0001, updateTol = 0.00001;
  int calls = 0;

  double energy = energyAt(e, phi);
  double energyNew = energy;
  int iter = 0;
  if (energy_trace && trace_cap > 0) energy_trace[0] = energyNew;
  if (max_iter <= 0) {
    if (final_energy) *final_energy = energy;
    if (n_armijo_calls) *n_armijo_calls = 0;
    free(de); free(dir); free(oldDir); free(oldGrad);
    return 0;
  }
  int stopping;
  do {
    iter++;
    energy = energyNew;
    /* calcTaudAndDescentDir */
    gradientAt(e, phi, de);
    for (int i = 0; i < n2; ++i) dir[i] = -de[i];
    for (int c = 0; c < 2; ++c) msor_vcycle(dir + c * NN, N, filterWidth, dir + c * NN);
    const double beta = ((iter - 1) % 10) ? (dot2(dir, dir, NN) / dot2(oldGrad, oldGrad, NN)) : 0.0;
    memcpy(oldGrad, dir, sizeof(double) * n2);
    for (int i = 0; i < n2; ++i) dir[i] += beta * oldDir[i];
    memcpy(oldDir, dir, sizeof(double) * n2);
    e->evals = 0;
    tau = armijo(e, phi, dir, de, tau, energy);
    if (tau_trace && calls < trace_cap) {
      tau_trace[calls] = tau;
      if (sigma_trace) sigma_trace[calls] = filterWidth;
      if (evals_trace) evals_trace[calls] = e->evals;
    }
    calls++;
    for (int i = 0; i < n2; ++i) dir[i] *= tau;
    if (tau == 0) {
      memcpy(dir, oldGrad, sizeof(double) * n2);
      memset(oldDir, 0, sizeof(double) * n2);
      e->evals = 0;
      tau = armijo(e, phi, dir, de, tau, energy);
      if (tau_trace && calls < trace_cap) {
        tau_trace[calls] = tau;
        if (sigma_trace) sigma_trace[calls] = filterWidth;
        if (evals_trace) evals_trace[calls] = e->evals;
      }
      calls++;
      for (int i = 0; i < n2; ++i) dir[i] *= tau;
    }
    if (tau != 0)
      for (int i = 0; i < n2; ++i) phi[i] += dir[i];
    energyNew = energyAt(e, phi);
    if (energy_trace && iter < trace_cap) energy_trace[iter] = energyNew;

    const double crit = grad_stop ? sqrt(dot2(de, de, NN)) : energy - energyNew; /* MultiVector::norm */
    stopping = (crit > stop_eps) && (iter < max_iter);
    if (!auto_sigma && tau == 0) stopping = 0;
    if (auto_sigma && (0.5 * filterWidth > stopFilterWidth) && ((energy - energyNew) < updateTol)) {
      filterWidth *= 0.5;
      tau = 1.;
      stopping = (filterWidth > stopFilterWidth) && (iter < max_iter);
    }
  } while (stopping);

  if (final_energy) *final_energy = energyNew;
  if (n_armijo_calls) *n_armijo_calls = calls;
  free(de); free(dir); free(oldDir); free(oldGrad);
  return iter;
}

int msor_level_solve(const double *R, const double *T, double *phi, int N, const msor_level_params *p,
                     double *energy_trace, double *tau_trace, double *sigma_trace, int *evals_trace,
                     int trace_cap, int *n_armijo_calls, double *final_energy) {
  const int NN = N * N, n2 = 2 * NN;
  Energy e = {R, T, N, p->lambda, (double *)malloc(sizeof(double) * n2), 0, p->data_term, NULL, NULL, NULL, 0,
              p->y_reg};
  if (p->data_term == MSOR_NCC) {
    e.Rn = (double *)malloc(sizeof(double) * NN);
    e.Tn = (double *)malloc(sizeof(double) * NN);
    msor_ncc_normalize(R, N, e.Rn);
  }
  const int iter = descent(&e, phi, p->max_iter, p->stop_eps, 1.0, 1, 0, energy_trace, tau_trace, sigma_trace,
                           evals_trace, trace_cap, n_armijo_calls, final_energy);
  free(e.tmp); free(e.Rn); free(e.Tn);
  return iter;
}

/* SeriesMatching::reduceDeformations (matchSeries.h:678-704): from phi = 0, minimise the sum of the
 * consistency energies of the n deformations with H1GradientDescent (fixed filter width sqrt(10 h),
 * gradientDescent.h:639-650), nonlinear CG and gradient-norm stopping. */
int msor_reduce_deformations(const double *psi, int n, int N, int max_iter, double stop_eps, double *phi,
                             double *final_energy) {
  const int n2 = 2 * N * N;
  Energy e = {NULL, NULL, N, 0.0, (double *)malloc(sizeof(double) * n2), 0, DT_CONSISTENCY, NULL, NULL, psi, n, 0.0};
  memset(phi, 0, sizeof(double) * n2);
  const int iter = descent(&e, phi, max_iter, stop_eps, sqrt(10 * gridH(N)), 0, 1, NULL, NULL, NULL, NULL, 0, NULL,
                           final_energy);
  free(e.tmp);
  return iter;
}

/* ---------------------------------------------------------------------------------------------
 * Input normalisation
 * ------------------------------------------------------------------------------------------- */

static int cmpd(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* prepareMultilevelArray (registration.h:331-347): saturated min/max (vec.h:864-870) + clamp +
 * scaleValuesTo01 (vec.h:1084-1089), or shift-to-zero and divide by the maximum. */
void msor_normalize(double *img, int count, double pct, int min_to_zero) {
  if (pct >= 0) {
    double *t = (double *)malloc(sizeof(double) * count);
    memcpy(t, img, sizeof(double) * count);
    qsort(t, count, sizeof(double), cmpd);
    const int m1 = count - 1;
    int i0 = (int)(pct / 200 * m1), i1 = (int)((1 - pct / 200) * m1);
    i0 = i0 < 0 ? 0 : (i0 > m1 ? m1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > m1 ? m1 : i1);
    const double lo = t[i0], hi = t[i1];
    free(t);
    for (int i = 0; i < count; ++i) img[i] = img[i] < lo ? lo : (img[i] > hi ? hi : img[i]);
    double mn = img[0];
    for (int i = 1; i < count; ++i) if (img[i] < mn) mn = img[i];
    for (int i = 0; i < count; ++i) img[i] += -1 * mn;
    double mx = img[0];
    for (int i = 1; i < count; ++i) if (img[i] > mx) mx = img[i];
    if (mx != 0)
      for (int i = 0; i < count; ++i) img[i] /= mx;
  } else {
    double mn = img[0];
    for (int i = 1; i < count; ++i) if (img[i] < mn) mn = img[i];
    if (mn < 0. || min_to_zero)
      for (int i = 0; i < count; ++i) img[i] += -mn;
    double mx = img[0];
    for (int i = 1; i < count; ++i) if (img[i] > mx) mx = img[i];
    if (mx != 0.)
      for (int i = 0; i < count; ++i) img[i] /= mx;
  }
}

/* ---------------------------------------------------------------------------------------------
 * Composition with harmonic extension
 * ------------------------------------------------------------------------------------------- */

/* DeformImage with an extension image of NaN (deformations.h:226-256) */
static void deformWithNaN(const double *img, const double *phi, int N, double *out) {
  const double h = gridH(N);
  const int NN = N * N;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      const int i = IDX(x, y, N);
      double ds[2];
      int inside = 1;
      ds[0] = x + phi[i] / h;
      ds[1] = y + phi[NN + i] / h;
      for (int c = 0; c < 2; ++c)
        if (ds[c] < 0.0 || ds[c] > (double)(N - 1)) inside = 0;
      out[i] = inside ? msor_interpolate(img, N, ds[0], ds[1]) : NAN;
    }
}

/* Stiffness matrix with Dirichlet rows/columns removed for defined nodes and unit diagonal there
 * (assembleAddMatrix with a Dirichlet mask and setDirichletNodes, FEOpInterface.h:624-639). */
static void dirichletStiffApply(const double *u, const unsigned char *defined, int N, double *out, double *tmp) {
  const int NN = N * N;
  for (int i = 0; i < NN; ++i) tmp[i] = defined[i] ? 0.0 : u[i];
  msor_stiff_apply(tmp, N, out);
  for (int i = 0; i < NN; ++i) out[i] = defined[i] ? u[i] : out[i];
}

static double dirichletDiag(int i, const unsigned char *defined, int N) {
  if (defined[i]) return 1.0;
  const int x = i % N, y = i / N;
  int ne = 0; /* number of adjacent elements */
  for (int ey = y - 1; ey <= y; ++ey)
    for (int ex = x - 1; ex <= x; ++ex)
      if (ex >= 0 && ey >= 0 && ex < N - 1 && ey < N - 1) ++ne;
  return ne * (4.0 / 6.0);
}

/* PCGInverse::apply (solver.h:264-359) with a diagonal preconditioner, absolute stopping at
 * 1e-16 on |g|^2, at most 1000 iterations; dest holds the start value. */
static void pcgDirichlet(const double *rhs, const unsigned char *defined, int N, double *dest) {
  const int NN = N * N;
  double *g = (double *)calloc(NN, sizeof(double));
  double *d = (double *)calloc(NN, sizeof(double));
  double *h = (double *)calloc(NN, sizeof(double));
  double *tmp = (double *)calloc(NN, sizeof(double));
  double *diag = (double *)malloc(sizeof(double) * NN);
  /* DiagonalPreconditioner stores the inverse |diagonal| (preconditioner.h:25-35) */
  for (int i = 0; i < NN; ++i) diag[i] = 1.0 / fabs(dirichletDiag(i, defined, N));
  dirichletStiffApply(dest, defined, N, h, tmp);
  for (int i = 0; i < NN; ++i) g[i] = h[i] - rhs[i];
  for (int i = 0; i < NN; ++i) h[i] = diag[i] * g[i];
  double spn = 0;
  for (int i = 0; i < NN; ++i) spn += g[i] * g[i];
  for (int i = 0; i < NN; ++i) d[i] -= h[i];
  int it = 0;
  while (!(spn < 1e-16) && it < 1000 && !isnan(spn)) {
    double an = 0;
    for (int i = 0; i < NN; ++i) an += g[i] * h[i];
    const double bd = an;
    dirichletStiffApply(d, defined, N, h, tmp);
    double ad = 0;
    for (int i = 0; i < NN; ++i) ad += d[i] * h[i];
    const double alpha = an / ad;
    for (int i = 0; i < NN; ++i) dest[i] += alpha * d[i];
    for (int i = 0; i < NN; ++i) g[i] += alpha * h[i];
    for (int i = 0; i < NN; ++i) h[i] = diag[i] * g[i];
    double bn = 0;
    for (int i = 0; i < NN; ++i) bn += g[i] * h[i];
    const double bf = bn / bd;
    for (int i = 0; i < NN; ++i) d[i] *= bf;
    for (int i = 0; i < NN; ++i) d[i] -= h[i];
    spn = 0;
    for (int i = 0; i < NN; ++i) spn += g[i] * g[i];
    ++it;
  }
  free(g); free(d); free(h); free(tmp); free(diag);
}

/* ConcatAndSmoothlyExtendDeformations (deformations.h:499-520) with SmoothlyExtendImage
 * (deformations.h:364-404). */
void msor_compose(const double *phi1, const double *phi2, int N, double *result) {
  const int NN = N * N;
  double *r = (double *)malloc(sizeof(double) * 2 * NN);
  for (int c = 0; c < 2; ++c) deformWithNaN(phi1 + c * NN, phi2, N, r + c * NN);
  for (int i = 0; i < 2 * NN; ++i) r[i] += phi2[i];
  unsigned char *defined = (unsigned char *)malloc(NN);
  for (int i = 0; i < NN; ++i) defined[i] = !isnan(r[i]);
  double *bc = (double *)malloc(sizeof(double) * NN);
  double *rhs = (double *)malloc(sizeof(double) * NN);
  for (int c = 0; c < 2; ++c) {
    for (int i = 0; i < NN; ++i) bc[i] = defined[i] ? r[c * NN + i] : 0.;
    msor_stiff_apply(bc, N, rhs);
    for (int i = 0; i < NN; ++i) rhs[i] *= -1.;
    for (int i = 0; i < NN; ++i) if (defined[i]) rhs[i] = bc[i];
    pcgDirichlet(rhs, defined, N, result + c * NN);
  }
  free(r); free(defined); free(bc); free(rhs);
}

/* ---------------------------------------------------------------------------------------------
 * Reconstruction statistics
 * ------------------------------------------------------------------------------------------- */

void msor_series_stats(const double *stack, int F, int NN, double fill, double *mean, double *count,
                       double *median) {
  double *s = (double *)malloc(sizeof(double) * (F > 0 ? F : 1));
  for (int j = 0; j < NN; ++j) {
    double m = 0;
    int c = 0;
    for (int f = 0; f < F; ++f) {
      const double v = stack[(size_t)f * NN + j];
      if (v != INFINITY) {
        m += v;
        s[c++] = v;
      }
    }
    if (c > 0) m /= c; else m = fill;
    mean[j] = m;
    count[j] = c;
    if (c == 0) {
      median[j] = fill;
    } else {
      qsort(s, c, sizeof(double), cmpd);
      median[j] = (c % 2 == 1) ? s[c / 2] : (s[c / 2 - 1] + s[c / 2]) / 2;
    }
  }
  free(s);
}

/* ---------------------------------------------------------------------------------------------
 * Series matching (one stage)
 * ------------------------------------------------------------------------------------------- */

typedef struct {
  int L;          /* finest level (precisionLevel) */
  double *lev[32]; /* images per level */
} Pyramid;

static void pyrBuild(Pyramid *p, const double *fine, int L) {
  p->L = L;
  for (int l = 0; l <= L; ++l) p->lev[l] = (double *)malloc(sizeof(double) * ((size_t)1 << (2 * l)));
  memcpy(p->lev[L], fine, sizeof(double) * ((size_t)1 << (2 * L)));
  for (int l = L; l > 0; --l) msor_restrict(p->lev[l], 1 << l, p->lev[l - 1]); /* MultilevelArray::levRestrict */
}
static void pyrFree(Pyramid *p) {
  for (int l = 0; l <= p->L; ++l) free(p->lev[l]);
}

/* deformation pyramid: 2 components per level */
typedef struct {
  int L;
  double *lev[32];
} DefPyramid;

static void defAlloc(DefPyramid *d, int L) {
  d->L = L;
  for (int l = 0; l <= L; ++l) d->lev[l] = (double *)calloc((size_t)2 << (2 * l), sizeof(double));
}
static void defFree(DefPyramid *d) {
  for (int l = 0; l <= d->L; ++l) free(d->lev[l]);
}
/* setTransformation + levRestrict (registration.h:549-553) */
static void defSet(DefPyramid *d, const double *fine) {
  const size_t NN = (size_t)1 << (2 * d->L);
  memcpy(d->lev[d->L], fine, sizeof(double) * 2 * NN);
  for (int l = d->L; l > 0; --l) {
    const size_t n = (size_t)1 << (2 * (l - 1));
    for (int c = 0; c < 2; ++c) msor_restrict(d->lev[l] + c * ((size_t)1 << (2 * l)), 1 << l, d->lev[l - 1] + c * n);
  }
}
/* RegistrationMultilevelDescentInterface::prolongate (registration.h:510-521) */
static void defProlong(DefPyramid *d, int l, int resample) {
  const int n = 1 << l, N = 2 * n;
  for (int c = 0; c < 2; ++c) {
    if (resample) msor_resample(d->lev[l] + c * n * n, n, d->lev[l + 1] + c * N * N, N);
    else msor_prolong(d->lev[l] + c * n * n, n, d->lev[l + 1] + c * N * N);
  }
}

static double levelLambda(const msor_series_params *p, int level) {
  const double f = p->lambda_factor > 0 ? pow(p->lambda_factor, p->stop_level - level) : 1; /* registration.h:523-526 */
  return f * p->lambda;
}

/* solveAndProlongToMaxDepth (registration.h:462-467) over MultilevelDescentInterface::solve
 * (quocDescent.h:83-105) with descentOnCurrentGrid (registration.h:1092-1112). */
static double solveML(const Pyramid *R, const Pyramid *T, DefPyramid *d, int start, int stop,
                      const msor_series_params *p) {
  double energy = NAN;
  for (int level = start; level <= stop; ++level) {
    msor_level_params lp = {levelLambda(p, level), p->max_iter, p->stop_eps, p->data_term, 0.0};
    msor_level_solve(R->lev[level], T->lev[level], d->lev[level], 1 << level, &lp, NULL, NULL, NULL, NULL, 0,
                     NULL, &energy);
    if (level < stop) defProlong(d, level, p->resample_prolong);
  }
  for (int l = stop; l < d->L; ++l) defProlong(d, l, p->resample_prolong);
  return energy;
}

/* In-place radix-2 transform of n = 2^k complex values (re, im interleaved), unnormalised,
 * exp(-2 pi i jk/n) forward and exp(+2 pi i jk/n) inverse; every twiddle from cos/sin directly. */
static void fft1(double *a, int n, int inverse) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = a[2 * i]; a[2 * i] = a[2 * j]; a[2 * j] = t;
      t = a[2 * i + 1]; a[2 * i + 1] = a[2 * j + 1]; a[2 * j + 1] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int k = 0; k < len / 2; ++k) {
      const double ang = 2.0 * 3.14159265358979323846 * k / len, wr = cos(ang), wi = inverse ? sin(ang) : -sin(ang);
      for (int i = 0; i < n; i += len) {
        double *u = a + 2 * (i + k), *v = a + 2 * (i + k + len / 2);
        const double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
        v[0] = u[0] - vr; v[1] = u[1] - vi;
        u[0] += vr; u[1] += vi;
      }
    }
  }
}

/* 2-D transform of an N x N complex array (x fastest): rows, then columns */
static void fft2(double *a, int N, int inverse) {
  double *col = (double *)malloc(sizeof(double) * 2 * N);
  for (int y = 0; y < N; ++y) fft1(a + (size_t)2 * y * N, N, inverse);
  for (int x = 0; x < N; ++x) {
    for (int y = 0; y < N; ++y) { col[2 * y] = a[2 * ((size_t)y * N + x)]; col[2 * y + 1] = a[2 * ((size_t)y * N + x) + 1]; }
    fft1(col, N, inverse);
    for (int y = 0; y < N; ++y) { a[2 * ((size_t)y * N + x)] = col[2 * y]; a[2 * ((size_t)y * N + x) + 1] = col[2 * y + 1]; }
  }
  free(col);
}

void msor_phase_correlation(const double *a, const double *b, int N, int normalise, double *out) {
  const size_t NN = (size_t)N * N;
  double *fa = (double *)calloc(2 * NN, sizeof(double)), *fb = (double *)calloc(2 * NN, sizeof(double));
  for (size_t i = 0; i < NN; ++i) { fa[2 * i] = a[i]; fb[2 * i] = b[i]; }
  fft2(fa, N, 0);
  fft2(fb, N, 0);
  for (size_t i = 0; i < NN; ++i) {
    /* FFT(a) * conj(FFT(b)), convolution.h:1095-1102 */
    const double tx = fa[2 * i] * fb[2 * i] + fa[2 * i + 1] * fb[2 * i + 1];
    const double ty = fa[2 * i + 1] * fb[2 * i] - fa[2 * i] * fb[2 * i + 1];
    const double nrm = normalise ? sqrt(tx * tx + ty * ty) : 1.0;
    const double scale = fabs(nrm) < 1e-15 ? 0.0 : 1.0 / nrm;
    fa[2 * i] = tx * scale;
    fa[2 * i + 1] = ty * scale;
  }
  fft2(fa, N, 1);
  for (size_t i = 0; i < NN; ++i) out[i] = fa[2 * i] * (1.0 / (double)NN);
  free(fa);
  free(fb);
}

void msor_register_translation(const double *reference, const double *templ, int N, int max_shift, int *shift) {
  const size_t NN = (size_t)N * N;
  double *pc = (double *)malloc(sizeof(double) * NN);
  msor_phase_correlation(reference, templ, N, max_shift <= 0, pc);
  int sx = 0, sy = 0;
  if (max_shift <= 0) { /* getMaxIndexAndValue: the first maximum */
    size_t at = 0;
    for (size_t i = 1; i < NN; ++i)
      if (pc[i] > pc[at]) at = i;
    sx = (int)(at % N);
    sy = (int)(at / N);
  } else { /* window |shift|_inf <= max_shift from (0, 0), x fastest, strict improvement */
    double best = pc[0];
    for (int dy = -max_shift; dy <= max_shift; ++dy)
      for (int dx = -max_shift; dx <= max_shift; ++dx) {
        const int px = (dx + N) % N, py = (dy + N) % N;
        if (pc[(size_t)py * N + px] > best) { best = pc[(size_t)py * N + px]; sx = px; sy = py; }
      }
  }
  shift[0] = sx <= N / 2 ? sx : sx - N;
  shift[1] = sy <= N / 2 ? sy : sy - N;
  free(pc);
}

/* TransformFunction<QC_2D>::interpolateTriangle (quoc/auxiliary.cpp:115-176): nodes of the bounding
 * box not yet set whose barycentric coordinates pass the 1e-6 tolerance take the affine
 * interpolation of the corner values (when finite) and are marked set. */
static void invTriangle(const double c[3][2], const double val[2][3], int N, unsigned char *set, double *dest) {
  int miny = N, maxy = -1, minx = N, maxx = -1;
  for (int k = 0; k < 3; ++k) {
    int qx = (int)c[k][0], qy = (int)c[k][1];
    if (qy > maxy) maxy = qy;
    if (qy < c[k][1]) ++qy;
    if (qy < miny) miny = qy;
    if (qx > maxx) maxx = qx;
    if (qx < c[k][0]) ++qx;
    if (qx < minx) minx = qx;
  }
  if (minx < 0) minx = 0;
  if (miny < 0) miny = 0;
  if (maxx > N - 1) maxx = N - 1;
  if (maxy > N - 1) maxy = N - 1;
  const double m00 = c[0][0] - c[2][0], m10 = c[0][1] - c[2][1];
  const double m01 = c[1][0] - c[2][0], m11 = c[1][1] - c[2][1];
  const double det = m00 * m11 - m10 * m01;
  if (det == 0 || !isfinite(det)) return;
  const double i00 = m11 / det, i01 = -m01 / det, i10 = -m10 / det, i11 = m00 / det;
  const size_t NN = (size_t)N * N;
  for (int x = minx; x <= maxx; ++x)
    for (int y = miny; y <= maxy; ++y) {
      const size_t at = (size_t)y * N + x;
      if (set[at]) continue;
      const double r0 = (double)x - c[2][0], r1 = (double)y - c[2][1];
      const double l0 = i00 * r0 + i01 * r1, l1 = i10 * r0 + i11 * r1;
      if (l0 < -1.e-6 || l1 < -1.e-6 || (l0 + l1) > 1. + 1.e-6) continue;
      set[at] = 1;
      for (int k = 0; k < 2; ++k) {
        const double v = l0 * val[k][0] + l1 * val[k][1] + (1.0 - l0 - l1) * val[k][2];
        if (isfinite(v)) dest[k * NN + at] = v;
      }
    }
}

void msor_approx_inverse(const double *phi, int N, double *inv) {
  const size_t NN = (size_t)N * N;
  const double H = 1.0 / (double)(N - 1);
  unsigned char *set = (unsigned char *)calloc(NN, 1);
  for (size_t i = 0; i < 2 * NN; ++i) inv[i] = 0.0; /* the extension image: zeros */
  for (int ey = 0; ey < N - 1; ++ey)
    for (int ex = 0; ex < N - 1; ++ex) {
      const int px[4] = {ex, ex + 1, ex, ex + 1}, py[4] = {ey, ey, ey + 1, ey + 1};
      double d[4][2], idv[2][4];
      for (int i = 0; i < 4; ++i) {
        const size_t at = (size_t)py[i] * N + px[i];
        double dx = (double)px[i] + phi[at] / H, dy = (double)py[i] + phi[NN + at] / H;
        dx = dx < (double)(N - 1) ? dx : (double)(N - 1); /* aol::Min, then aol::Max */
        dy = dy < (double)(N - 1) ? dy : (double)(N - 1);
        dx = dx < 0.0 ? 0.0 : dx;
        dy = dy < 0.0 ? 0.0 : dy;
        d[i][0] = dx;
        d[i][1] = dy;
        idv[0][i] = 1.0 * ((double)px[i] * H - 0.0); /* generateIdentity */
        idv[1][i] = 1.0 * ((double)py[i] * H - 0.0);
      }
      double c[3][2] = {{d[0][0], d[0][1]}, {d[1][0], d[1][1]}, {d[2][0], d[2][1]}};
      double val[2][3] = {{idv[0][0], idv[0][1], idv[0][2]}, {idv[1][0], idv[1][1], idv[1][2]}};
      invTriangle(c, val, N, set, inv);
      c[0][0] = d[3][0];
      c[0][1] = d[3][1];
      val[0][0] = idv[0][3];
      val[1][0] = idv[1][3];
      invTriangle(c, val, N, set, inv);
    }
  for (int k = 0; k < 2; ++k) {
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (size_t i = 0; i < NN; ++i) {
      const int y = (int)(i / N), x = (int)(i % N);
      inv[k * NN + i] -= 1.0 * ((double)(k == 0 ? x : y) * H - 0.0);
      if (set[i]) {
        if (inv[k * NN + i] < mn) mn = inv[k * NN + i];
        if (inv[k * NN + i] > mx) mx = inv[k * NN + i];
      }
    }
    const double fill = 0.5 * (mn + mx);
    for (size_t i = 0; i < NN; ++i)
      if (!set[i]) inv[k * NN + i] = fill;
  }
  free(set);
}

static void prepared(const double *raw, int NN, const msor_series_params *p, double *out) {
  memcpy(out, raw, sizeof(double) * NN);
  if (!p->dont_normalize) msor_normalize(out, NN, p->saturation_pct, p->normalize_min_to_zero);
}

int msor_match_series(const double *templates, int n, const double *reference, int N,
                      const msor_series_params *p, const double *pairwise_in,
                      double *deformations, double *pairwise_out, double *energies,
                      double *mean, double *count, double *median, double *reduced) {
  int L = 0;
  while ((1 << L) < N) ++L;
  const size_t NN = (size_t)N * N;
  double *img = (double *)malloc(sizeof(double) * NN);
  double *acc = (double *)calloc(2 * NN, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * 2 * NN);
  Pyramid ref0, ref, tpl;
  prepared(reference, (int)NN, p, img);
  pyrBuild(&ref0, img, L);
  pyrBuild(&ref, img, L);
  DefPyramid def;
  defAlloc(&def, L);

  for (int i = 0; i < n; ++i) {
    double *zero = (double *)calloc(2 * NN, sizeof(double));
    defSet(&def, zero); /* setTransformationToZero */
    free(zero);
    if (i == 0 && p->pair0_init) defSet(&def, p->pair0_init); /* reduceDeformations, stage > 1 (L123-135) */
    prepared(templates + i * NN, (int)NN, p, img);
    pyrBuild(&tpl, img, L);
    if (p->corr_init && (pairwise_in == NULL || i == 0)) {
      /* useCorrelationToInitTranslation (matchSeries.h:145-153): registerImages(template,
       * reference) in pixels, times h, on every level (setTransformationToTranslation) */
      int sh[2];
      msor_register_translation(tpl.lev[L], ref.lev[L], N, p->max_corr_shift, sh);
      const double h = 1.0 / (double)(N - 1);
      for (size_t k = 0; k < NN; ++k) { tmp[k] = sh[0] * h; tmp[NN + k] = sh[1] * h; }
      defSet(&def, tmp);
    }
    if (pairwise_in == NULL || i == 0) {
      energies[i] = solveML(&ref, &tpl, &def, p->start_level, p->stop_level, p);
      if (p->first_try_out) memcpy(p->first_try_out + i * 2 * NN, def.lev[L], sizeof(double) * 2 * NN);
      if (p->alt_start_level > 0 && p->alt_start_level != p->start_level) {
        /* altStartLevel (matchSeries.h:158-203): a second solve from zero at altStartLevel; the first
         * is kept when it leaves fewer template pixels outside the domain (or, with
         * energyAsAltCriterion, has the lower energy) */
        double *first = (double *)malloc(sizeof(double) * 2 * NN), *warped = (double *)malloc(sizeof(double) * NN);
        memcpy(first, def.lev[L], sizeof(double) * 2 * NN);
        const double e1 = energies[i];
        msor_deform_image(tpl.lev[L], first, N, INFINITY, 0, warped);
        long out1 = 0, out2 = 0;
        for (size_t k = 0; k < NN; ++k) out1 += isinf(warped[k]) && warped[k] > 0;
        double *zero = (double *)calloc(2 * NN, sizeof(double));
        defSet(&def, zero);
        free(zero);
        const double e2 = solveML(&ref, &tpl, &def, p->alt_start_level, p->stop_level, p);
        msor_deform_image(tpl.lev[L], def.lev[L], N, INFINITY, 0, warped);
        for (size_t k = 0; k < NN; ++k) out2 += isinf(warped[k]) && warped[k] > 0;
        const int keepFirst = p->energy_as_alt_criterion ? (e1 < e2) : (out1 < out2);
        if (keepFirst) defSet(&def, first);
        energies[i] = keepFirst ? e1 : e2;
        free(first);
        free(warped);
      }
    } else {
      defSet(&def, pairwise_in + i * 2 * NN); /* loadTransformation of the previous stage */
    }
    if (pairwise_out) memcpy(pairwise_out + i * 2 * NN, def.lev[L], sizeof(double) * 2 * NN);
    if (!p->dont_accumulate) {
      memcpy(tmp, acc, sizeof(double) * 2 * NN);
      msor_compose(def.lev[L], tmp, N, acc); /* composeDeformations(current, old acc) -> acc */
      if (i > 0) {
        defSet(&def, acc);
        pyrFree(&ref);
        pyrBuild(&ref, ref0.lev[L], L); /* setReference(reference) */
        energies[i] = solveML(&ref, &tpl, &def, p->refine_start_level, p->refine_stop_level, p);
        memcpy(acc, def.lev[L], sizeof(double) * 2 * NN);
      }
      pyrFree(&ref);
      pyrBuild(&ref, tpl.lev[L], L); /* setReference(template) */
    }
    memcpy(deformations + i * 2 * NN, (!p->dont_accumulate && i > 0) ? acc : def.lev[L], sizeof(double) * 2 * NN);
    if (!p->dont_accumulate && p->calc_inverse) {
      /* calcInverseDeformation (matchSeries.h:247-259): the template as reference, the reference
       * as template, from the approximate inverse of the accumulated deformation, refine levels */
      double *inv0 = (double *)malloc(sizeof(double) * 2 * NN);
      msor_approx_inverse(acc, N, inv0);
      defSet(&def, inv0);
      free(inv0);
      solveML(&tpl, &ref0, &def, p->refine_start_level, p->refine_stop_level, p);
      if (p->inverse_out) memcpy(p->inverse_out + i * 2 * NN, def.lev[L], sizeof(double) * 2 * NN);
    }
    pyrFree(&tpl);
  }

  /* reduceDeformations (L565-570, 678-704), then averageSeries (matchSeries.h:353-451): raw templates
   * (first deformed by the reduced deformation), Inf outside, fill = last template mean */
  double *red = NULL, *pre = NULL;
  if (p->reduce_deformations && n > 0) {
    red = (double *)malloc(sizeof(double) * 2 * NN);
    pre = (double *)malloc(sizeof(double) * NN);
    msor_reduce_deformations(deformations, n, N, p->reduce_max_iter, p->reduce_stop_eps, red, NULL);
    if (reduced) memcpy(reduced, red, sizeof(double) * 2 * NN);
  }
  double *stack = (double *)malloc(sizeof(double) * NN * (n > 0 ? n : 1));
  double lastMean = 0;
  for (int i = 0; i < n; ++i) {
    if (red) {
      msor_deform_image(templates + i * NN, red, N, INFINITY, 0, pre);
      msor_deform_image(pre, deformations + i * 2 * NN, N, INFINITY, 0, stack + i * NN);
    } else {
      msor_deform_image(templates + i * NN, deformations + i * 2 * NN, N, INFINITY, 0, stack + i * NN);
    }
    if (i == n - 1) {
      double s = 0;
      for (size_t j = 0; j < NN; ++j) s += templates[i * NN + j];
      lastMean = s / NN;
    }
  }
  msor_series_stats(stack, n, (int)NN, lastMean, mean, count, median);
  free(stack);
  free(red);
  free(pre);
  pyrFree(&ref0);
  pyrFree(&ref);
  defFree(&def);
  free(img);
  free(acc);
  free(tmp);
  return 0;
}
