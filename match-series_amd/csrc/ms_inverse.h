// Approximate inverse of a deformation on the GPU (qc::approxInverseDeformation,
// quoc/deformations.h:341-355, over TransformFunction<QC_2D>::transform and interpolateTriangle,
// quoc/auxiliary.cpp:33-176). The reference rasterises the deformed grid element by element (x
// fastest), two triangles per element, and a node takes the value of the first triangle that covers
// it. Here every triangle marks the nodes it covers with atomicMin of its sequence number
// (2 * element + triangle), so each node ends with the first one in the reference's order; a second
// pass interpolates from that triangle. FP64 arithmetic in the reference's operation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

namespace msg {

constexpr unsigned INV_NONE = 0xffffffffu;

// The four corners of element (ex, ey) moved by phi, in grid units, clamped to [0, n-1]
// (aol::Min, then aol::Max), in the order (x, y), (x+1, y), (x, y+1), (x+1, y+1).
template <typename S>
__device__ inline void inv_corners(const S *__restrict__ phi, int n, double H, int ex, int ey, double d[4][2]) {
  const size_t nn = (size_t)n * n;
  const double top = (double)(n - 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int px = ex + (i & 1), py = ey + (i >> 1);
    const size_t at = (size_t)py * n + px;
    double dx = (double)px + (double)phi[at] / H, dy = (double)py + (double)phi[nn + at] / H;
    dx = dx < top ? dx : top;
    dy = dy < top ? dy : top;
    dx = dx < 0.0 ? 0.0 : dx;
    dy = dy < 0.0 ? 0.0 : dy;
    d[i][0] = dx;
    d[i][1] = dy;
  }
}

// Triangle t of an element: corners (0, 1, 2) for t = 0, (3, 1, 2) for t = 1. Its node bounding box,
// the inverse of [c0 - c2, c1 - c2] (Matrix22::make_inverse: adjugate divided by the determinant);
// ok = false for a zero or non-finite determinant.
struct InvTri {
  double c2x, c2y, i00, i01, i10, i11;
  int minx, maxx, miny, maxy;
  bool ok;
};

__device__ inline InvTri inv_triangle(const double d[4][2], int t, int n) {
  const double c[3][2] = {{d[t ? 3 : 0][0], d[t ? 3 : 0][1]}, {d[1][0], d[1][1]}, {d[2][0], d[2][1]}};
  InvTri r;
  r.minx = n;
  r.miny = n;
  r.maxx = -1;
  r.maxy = -1;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int qx = (int)c[k][0], qy = (int)c[k][1];
    r.maxy = qy > r.maxy ? qy : r.maxy;
    if (qy < c[k][1]) ++qy;
    r.miny = qy < r.miny ? qy : r.miny;
    r.maxx = qx > r.maxx ? qx : r.maxx;
    if (qx < c[k][0]) ++qx;
    r.minx = qx < r.minx ? qx : r.minx;
  }
  r.minx = r.minx > 0 ? r.minx : 0;
  r.miny = r.miny > 0 ? r.miny : 0;
  r.maxx = r.maxx < n - 1 ? r.maxx : n - 1;
  r.maxy = r.maxy < n - 1 ? r.maxy : n - 1;
  const double m00 = c[0][0] - c[2][0], m10 = c[0][1] - c[2][1];
  const double m01 = c[1][0] - c[2][0], m11 = c[1][1] - c[2][1];
  const double det = m00 * m11 - m10 * m01;
  r.ok = det != 0 && isfinite(det);
  r.i00 = m11 / det;
  r.i01 = -m01 / det;
  r.i10 = -m10 / det;
  r.i11 = m00 / det;
  r.c2x = c[2][0];
  r.c2y = c[2][1];
  return r;
}

// Barycentric coordinates of node (x, y); true when inside up to the reference's 1e-6 tolerance
__device__ inline bool inv_inside(const InvTri &t, int x, int y, double &l0, double &l1) {
  const double r0 = (double)x - t.c2x, r1 = (double)y - t.c2y;
  l0 = t.i00 * r0 + t.i01 * r1;
  l1 = t.i10 * r0 + t.i11 * r1;
  return !(l0 < -1.e-6 || l1 < -1.e-6 || (l0 + l1) > 1. + 1.e-6);
}

__global__ void k_fill_u32(unsigned *__restrict__ p, size_t cnt, unsigned v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// One thread per element: every node its two triangles cover takes the smaller sequence number.
template <typename S>
__global__ void k_inv_keys(const S *__restrict__ phi, int n, double H, unsigned *__restrict__ key) {
  const long long ne = (long long)(n - 1) * (n - 1);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int ex = (int)(e % (n - 1)), ey = (int)(e / (n - 1));
  double d[4][2];
  inv_corners(phi, n, H, ex, ey, d);
  for (int t = 0; t < 2; ++t) {
    const InvTri tri = inv_triangle(d, t, n);
    if (!tri.ok) continue;
    const unsigned seq = (unsigned)(2 * e + t);
    for (int x = tri.minx; x <= tri.maxx; ++x)
      for (int y = tri.miny; y <= tri.maxy; ++y) {
        double l0, l1;
        if (inv_inside(tri, x, y, l0, l1)) atomicMin(&key[(size_t)y * n + x], seq);
      }
  }
}

// One thread per node: the value interpolated from its triangle (the zero extension where that is
// not finite), minus the identity; nodes without a triangle are left for k_inv_fill.
template <typename S>
__global__ void k_inv_values(const S *__restrict__ phi, int n, double H, const unsigned *__restrict__ key,
                             S *__restrict__ out) {
  const size_t nn = (size_t)n * n;
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  const unsigned k = key[p];
  if (k == INV_NONE) return;
  const int x = (int)(p % n), y = (int)(p / n);
  const long long e = k >> 1;
  const int t = (int)(k & 1);
  const int ex = (int)(e % (n - 1)), ey = (int)(e / (n - 1));
  double d[4][2];
  inv_corners(phi, n, H, ex, ey, d);
  const InvTri tri = inv_triangle(d, t, n);
  double l0, l1;
  inv_inside(tri, x, y, l0, l1);
  // corner values: the identity (generateIdentity: 1 * (node * H - 0)) at corners (t ? 3 : 0), 1, 2
  const int cx0 = ex + (t ? 1 : 0), cy0 = ey + (t ? 1 : 0);
  const double vx0 = 1.0 * ((double)cx0 * H - 0.0), vy0 = 1.0 * ((double)cy0 * H - 0.0);
  const double vx1 = 1.0 * ((double)(ex + 1) * H - 0.0), vy1 = 1.0 * ((double)ey * H - 0.0);
  const double vx2 = 1.0 * ((double)ex * H - 0.0), vy2 = 1.0 * ((double)(ey + 1) * H - 0.0);
  const double w2 = 1.0 - l0 - l1;
  double vx = l0 * vx0 + l1 * vx1 + w2 * vx2, vy = l0 * vy0 + l1 * vy1 + w2 * vy2;
  vx = isfinite(vx) ? vx : 0.0;
  vy = isfinite(vy) ? vy : 0.0;
  out[p] = (S)(vx - 1.0 * ((double)x * H - 0.0));
  out[nn + p] = (S)(vy - 1.0 * ((double)y * H - 0.0));
}

// Per workgroup of 256: min and max of both components over the covered nodes
// (getMinValueMasked / getMaxValueMasked start from +-DBL_MAX).
template <typename S>
__global__ __launch_bounds__(256) void k_inv_minmax(const S *__restrict__ out, const unsigned *__restrict__ key,
                                                    size_t nn, double *__restrict__ part) {
  __shared__ double s[4][256];
  double mnx = DBL_MAX, mxx = -DBL_MAX, mny = DBL_MAX, mxy = -DBL_MAX;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
    if (key[i] == INV_NONE) continue;
    const double a = (double)out[i], b = (double)out[nn + i];
    mnx = a < mnx ? a : mnx;
    mxx = a > mxx ? a : mxx;
    mny = b < mny ? b : mny;
    mxy = b > mxy ? b : mxy;
  }
  s[0][threadIdx.x] = mnx;
  s[1][threadIdx.x] = mxx;
  s[2][threadIdx.x] = mny;
  s[3][threadIdx.x] = mxy;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const int j = threadIdx.x + o;
      s[0][threadIdx.x] = s[0][j] < s[0][threadIdx.x] ? s[0][j] : s[0][threadIdx.x];
      s[1][threadIdx.x] = s[1][j] > s[1][threadIdx.x] ? s[1][j] : s[1][threadIdx.x];
      s[2][threadIdx.x] = s[2][j] < s[2][threadIdx.x] ? s[2][j] : s[2][threadIdx.x];
      s[3][threadIdx.x] = s[3][j] > s[3][threadIdx.x] ? s[3][j] : s[3][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// Nodes no triangle covers: the midpoint of the covered values' range, per component
template <typename S>
__global__ void k_inv_fill(S *__restrict__ out, const unsigned *__restrict__ key, size_t nn, double fx, double fy) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x)
    if (key[i] == INV_NONE) {
      out[i] = (S)fx;
      out[nn + i] = (S)fy;
    }
}

}  // namespace msg
