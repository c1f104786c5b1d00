// FP32-arithmetic kernels for FP32 fields: the SSD energy pass, the SSD variation and the V-cycle
// of the H1 metric. Same launch geometry, partial-sum layout and pair-state protocol as the FP64
// kernels in ms_kernels.h, which FP64 fields keep (they reproduce the reference's arithmetic).
//
// Numerics: the warp and the stencils run in FP32 with fused multiply-adds; energy sums, slopes and
// the step decisions stay FP64 (per-element values enter FP64 accumulators). A deformed quad point
// is computed in grid units relative to its element, e + u with u = quad offset + displacement / h,
// so the local coordinate keeps FP32 precision of the displacement (about 1e-6 px at 10 px) rather
// than of the absolute position (about 6e-5 px at 1023). Division by h becomes a multiplication by
// n - 1 and the Jacobi division by the diagonal a multiplication by its reciprocal.
#pragma once

#include "ms_kernels.h"

namespace msg {

__device__ __forceinline__ float qpxf(int q) { return (q >> 1) ? (float)MSG_QP1 : (float)MSG_QP0; }
__device__ __forceinline__ float qpyf(int q) { return (q & 1) ? (float)MSG_QP1 : (float)MSG_QP0; }
// basis value of dof b at quad point q (compile-time constants once the loops are unrolled)
__device__ __forceinline__ float bqf(int b, int q) { return (float)bq(b, q); }

// One axis of a deformed quad point: the position e + u in grid units (e the element's lower node)
// clipped to [0, n-1] as transformAndClipCoord does; `in` is false on a clipped axis. The element
// index stays in [0, n-2] for every input (NaN included).
__device__ __forceinline__ void clip_rel(int e, float u, int n, int &el, float &loc, bool &in) {
  const float lo = -(float)e, hi = (float)(n - 1 - e);
  in = (u >= lo) && (u < hi);
  const float uc = fmaxf(u, lo);  // NaN -> lo
  const float fl = floorf(uc);
  if (uc < hi) {
    el = e + (int)fl;
    loc = uc - fl;
  } else {
    el = n - 2;
    loc = 1.f;
  }
}

struct HitF {
  int i0;  // lower-left node of the element hit
  float lx, ly;
  bool inx, iny;
};

// quad point of element (ex, ey) at ux, uy grid units from the element's lower node
__device__ __forceinline__ HitF hit_rel(int ex, int ey, float ux, float uy, int n) {
  HitF r;
  int elx, ely;
  clip_rel(ex, ux, n, elx, r.lx, r.inx);
  clip_rel(ey, uy, n, ely, r.ly, r.iny);
  r.i0 = elx + ely * n;
  return r;
}

// Values of the Q1 interpolant of the element's nodal values p at the 4 quad points (q = 2 qx + qy),
// interpolating along x, then y.
__device__ __forceinline__ void quad_values(const float p[4], float v[4]) {
#pragma clang fp contract(fast)
  const float q0 = (float)MSG_QP0, q1 = (float)MSG_QP1;
  const float dx0 = p[1] - p[0], dx1 = p[3] - p[2];
  const float a0 = q0 * dx0 + p[0], a1 = q1 * dx0 + p[0];  // row y = 0 at qx = QP0, QP1
  const float b0 = q0 * dx1 + p[2], b1 = q1 * dx1 + p[2];  // row y = 1
  v[0] = q0 * (b0 - a0) + a0;
  v[1] = q1 * (b0 - a0) + a0;
  v[2] = q0 * (b1 - a1) + a1;
  v[3] = q1 * (b1 - a1) + a1;
}

// Gathers go through a buffer resource over one image: 32-bit offsets (no 64-bit address arithmetic
// per load), neighbouring nodes merged into one two-dword load, and reads past the image return 0.
// Resource word 3 is the gfx9-family raw-buffer format (32-bit data, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t image_rsrc(const float *p, int count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, count * 4, 0x00020000);
}
__device__ __forceinline__ float ld_rsrc(__amdgpu_buffer_rsrc_t r, unsigned idx) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4u, 0, 0));
}

// bilinear value at the hit; t receives the 4 nodal values
__device__ __forceinline__ float bilin_f(__amdgpu_buffer_rsrc_t T, const HitF &p, int n, float t[4]) {
#pragma clang fp contract(fast)
  const unsigned o = (unsigned)p.i0, on = o + (unsigned)n;
  t[0] = ld_rsrc(T, o);
  t[1] = ld_rsrc(T, o + 1u);
  t[2] = ld_rsrc(T, on);
  t[3] = ld_rsrc(T, on + 1u);
  const float a = p.lx * (t[1] - t[0]) + t[0];
  const float b = p.lx * (t[3] - t[2]) + t[2];
  return p.ly * (b - a) + a;
}

// a^T K_loc b of the Q1 stiffness matrix in the difference form of stiff_form, as its edge and
// diagonal sums e and d in FP32; the caller applies their weights 1/6 and 2/6 in FP64 (FP32's 1/6 is
// 3e-8 too large, a bias that every element would share)
__device__ __forceinline__ void stiff_parts_f(const float a[4], const float b[4], float &e, float &d) {
#pragma clang fp contract(fast)
  const float ea0 = a[0] - a[1], ea1 = a[0] - a[2], ea2 = a[1] - a[3], ea3 = a[2] - a[3];
  const float da0 = a[0] - a[3], da1 = a[1] - a[2];
  const float eb0 = b[0] - b[1], eb1 = b[0] - b[2], eb2 = b[1] - b[3], eb3 = b[2] - b[3];
  const float db0 = b[0] - b[3], db1 = b[1] - b[2];
  e = ea0 * eb0 + ea1 * eb1 + ea2 * eb2 + ea3 * eb3;
  d = da0 * db0 + da1 * db1;
}

// One axis of a deformed quad point for the energy's value: the position e + u clamped to the grid,
// [0, n-1] in grid units (lo = -e, hi = n-1-e), as element offset `off` from e and local coordinate
// `loc`. At the upper bound the element is the last node's (offset hi, loc 0) and the bilinear form
// returns that node's value exactly -- the interpolant's value there, which the reference's clip to
// width - eps reaches to rounding; the variation keeps clip_rel, which also reports a clipped axis.
// NaN clamps to the lower bound, so the node index stays in [0, n-1].
__device__ __forceinline__ void clamp_rel(float u, float lo, float hi, int &off, float &loc) {
  const float uc = fminf(fmaxf(u, lo), hi);
  const float fl = floorf(uc);
  off = (int)fl;
  loc = uc - fl;
}

// bilinear value of T at byte offset vo + (lx, ly) (reads past the image return 0 and meet a weight of
// 0); the upper row's loads take the row pitch as their scalar offset
__device__ __forceinline__ float bilin_at(__amdgpu_buffer_rsrc_t T, unsigned vo, int n, float lx, float ly) {
#pragma clang fp contract(fast)
  const unsigned pitch = (unsigned)n * 4u;
  auto ld = [&](unsigned extra, unsigned so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(T, vo + extra, so, 0));
  };
  const float t0 = ld(0u, 0u), t1 = ld(4u, 0u), t2 = ld(0u, pitch), t3 = ld(4u, pitch);
  const float a = lx * (t1 - t0) + t0;
  const float b = lx * (t3 - t2) + t2;
  return ly * (b - a) + a;
}

// ================================================================================================
// SSD energy pass for up to PLAN_CAND step sizes (k_energy in FP32 arithmetic). Tiles of 64 x 4 RUN
// elements; each thread takes a vertical run of RUN elements, reusing the upper node row of one
// element as the lower row of the next.
// ================================================================================================
// FIDX: byte offsets of the gathers formed in FP32 arithmetic, exact while n^2 <= 2^24 (n <= 4096):
// two fused multiply-adds and one conversion instead of two conversions and integer arithmetic.
template <bool FIDX, int RUN>
__global__ void __launch_bounds__(256)
k_energy_f32(const float *__restrict__ R, const float *__restrict__ T, const float *__restrict__ phi,
             const float *__restrict__ z, const float *__restrict__ dold, int n, double vol,
             const PairState *st, double *__restrict__ part, int ntx, int nty, int npairs) {
#pragma clang fp contract(fast)
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  const int nc = st[pair].ncand;
  if (nc == 0 || st[pair].exact) return;  // uniform across the workgroup (k_energy_exact_f32 takes exact pairs)
  const int dm = st[pair].dirmode;
  // Passes evaluate at most PLAN_CAND step sizes (armijo_plan; 2 for an FP64 re-evaluation, 1 for a
  // plain energy): candidate slots beyond it are never filled.
  constexpr int NC = PLAN_CAND;
  constexpr int NACC = NC + 3;  // E_SSD per candidate; sums of f'Kf, f'Kd, d'Kd
  __shared__ double red[4 * NACC];

  const size_t nn = (size_t)n * n;
  const float *Rp = R + pair * nn, *Tp = T + pair * nn;
  const __amdgpu_buffer_rsrc_t Trs = image_rsrc(Tp, (int)nn);
  const float *px = phi + pair * 2 * nn, *py = px + nn;
  const float *zx = z + pair * 2 * nn, *zy = zx + nn, *ox_ = dold + pair * 2 * nn, *oy_ = ox_ + nn;
  const bool cg = (dm == DIR_CG), use_old = st[pair].cg_old != 0;
  const double beta = st[pair].beta;
  const bool use_dir = (dm != DIR_NONE);
  const float nm1 = (float)(n - 1);
  float tau[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) tau[k] = (k < nc) ? (float)st[pair].cand[k] : 0.f;  // powers of two: exact

  const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
  const int ey0 = (tile / ntx) * (4 * RUN) + (threadIdx.x / EN_TX) * RUN;

  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) acc[k] = 0.0;

  if (ex < n - 1) {
    // A run's sums in FP32 (a few terms each), converted once: per step size the SSD sum, and the
    // e and d parts of f'Kf, f'Kd, d'Kd
    float ssd[NC], se[3] = {0.f, 0.f, 0.f}, sd[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NC; ++k) ssd[k] = 0.f;
    // node values of the element, b = x + 2 y: slots 0, 1 the lower row, 2, 3 the upper row
    float f0[4], f1[4], d0[4], d1[4], rn[4];
    auto load_row = [&](int y, int slot) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int g = ex + c + y * n;
        f0[slot + c] = px[g];
        f1[slot + c] = py[g];
        d0[slot + c] = use_dir ? search_dir<float>(zx, ox_, g, cg, use_old, beta) : 0.f;
        d1[slot + c] = use_dir ? search_dir<float>(zy, oy_, g, cg, use_old, beta) : 0.f;
        rn[slot + c] = Rp[g];
      }
    };
    if (ey0 < n - 1) load_row(ey0, 0);
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
      const int ey = ey0 + j;
      if (ey >= n - 1) break;  // uniform across a wave (a wave is one row of 64 elements)
      if (j > 0) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          f0[c] = f0[2 + c];
          f1[c] = f1[2 + c];
          d0[c] = d0[2 + c];
          d1[c] = d1[2 + c];
          rn[c] = rn[2 + c];
        }
      }
      load_row(ey + 1, 2);
      // Positions are linear in the step size: the quad points of phi + tau d sit at u_f + tau u_d
      // (grid units relative to the element), and psi^T K psi of phi + tau d is a quadratic in tau. One
      // interpolation of phi and d per element serves every candidate; positions differ from those of
      // the stored phi + tau d (rounded per node) by that rounding, about 3e-7 px at 10 px.
      float rq[4], ufx[4], ufy[4], udx[4], udy[4];
      quad_values(rn, rq);
      quad_values(f0, ufx);
      quad_values(f1, ufy);
      quad_values(d0, udx);
      quad_values(d1, udy);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ufx[q] = ufx[q] * nm1 + qpxf(q);
        ufy[q] = ufy[q] * nm1 + qpyf(q);
        udx[q] *= nm1;
        udy[q] *= nm1;
      }
      {
        float e, d;
        stiff_parts_f(f0, f0, e, d);
        se[0] += e;
        sd[0] += d;
        stiff_parts_f(f1, f1, e, d);
        se[0] += e;
        sd[0] += d;
        stiff_parts_f(f0, d0, e, d);
        se[1] += e;
        sd[1] += d;
        stiff_parts_f(f1, d1, e, d);
        se[1] += e;
        sd[1] += d;
        stiff_parts_f(d0, d0, e, d);
        se[2] += e;
        sd[2] += d;
        stiff_parts_f(d1, d1, e, d);
        se[2] += e;
        sd[2] += d;
      }
      // the quad point clamped to [0, n-1] in grid units: lo = -e, hi = n-1-e relative to the element
      const float lox = -(float)ex, hix = (float)(n - 1 - ex), loy = -(float)ey, hiy = (float)(n - 1 - ey);
      const int gbase = ex + ey * n;
      const float base4 = 4.f * (float)gbase, pitch4 = 4.f * (float)n;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        if (k < nc) {
          float a = 0.f;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            float lx, ly;
            unsigned vo;
            if constexpr (FIDX) {
              // med3 clamps; a NaN position gives offset 0 (the conversion's NaN result), in bounds
              const float cx = __builtin_amdgcn_fmed3f(tau[k] * udx[q] + ufx[q], lox, hix);
              const float cy = __builtin_amdgcn_fmed3f(tau[k] * udy[q] + ufy[q], loy, hiy);
              const float fx = floorf(cx), fy = floorf(cy);
              lx = cx - fx;
              ly = cy - fy;
              // 4 (gbase + fx + fy n): integers below 2^26 that are multiples of 4, exact in FP32
              vo = (unsigned)(fy * pitch4 + (fx * 4.f + base4));
            } else {
              int ox, oy;
              clamp_rel(tau[k] * udx[q] + ufx[q], lox, hix, ox, lx);
              clamp_rel(tau[k] * udy[q] + ufy[q], loy, hiy, oy, ly);
              vo = (unsigned)(gbase + ox + oy * n) * 4u;
            }
            const float t = bilin_at(Trs, vo, n, lx, ly) - rq[q];
            a += t * t;
          }
          ssd[k] += a;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = (double)ssd[k] * (0.125 * vol);  // 1/2 t^2 * quad weight 1/4 * vol
#pragma unroll
    for (int m = 0; m < 3; ++m) acc[NC + m] = (double)se[m] * (1.0 / 6.0) + (double)sd[m] * (2.0 / 6.0);
  }
  const int nblk = tiles;
  block_sum_n<NACC>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (k < nc) {
        const double taud = st[pair].cand[k];
        part[((size_t)(pair * MAX_CAND + k) * 2 + 0) * nblk + tile] = acc[k];
        part[((size_t)(pair * MAX_CAND + k) * 2 + 1) * nblk + tile] =
            0.5 * (acc[NC] + taud * (2.0 * acc[NC + 1] + taud * acc[NC + 2]));
      }
  }
}

// The same pass in FP64 arithmetic at the stored positions (energy_elem<float>, the pass of FP64 data
// terms on FP32 fields) for the pairs whose control flagged a near-tie step decision (`exact`), on the
// tiles and partial layout of k_energy_f32. Its own kernel so that the FP32 pass keeps its register
// budget; a grid-stride loop over the tiles of the flagged pairs, so that a pass without any costs one
// check per pair.
__global__ void __launch_bounds__(256)
k_energy_exact_f32(const float *__restrict__ R, const float *__restrict__ T, const float *__restrict__ phi,
                   const float *__restrict__ z, const float *__restrict__ dold, int n, double h, int width,
                   const PairState *st, double *__restrict__ part, int ntx, int nty, int npairs) {
  const int tiles = ntx * nty;
  const size_t nn = (size_t)n * n;
  const int run = en32_run(n);
  for (int pair = 0; pair < npairs; ++pair) {
    if (!st[pair].exact || st[pair].ncand == 0) continue;  // uniform across the workgroup
    const float *Rp = R + pair * nn, *Tp = T + pair * nn;
    const float *px = phi + pair * 2 * nn, *py = px + nn;
    const float *zx = z + pair * 2 * nn, *zy = zx + nn, *ox_ = dold + pair * 2 * nn, *oy_ = ox_ + nn;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
      const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
      const int ey0 = (tile / ntx) * (4 * run) + (threadIdx.x / EN_TX) * run;
      double accS[MAX_CAND], accR[MAX_CAND];
#pragma unroll
      for (int k = 0; k < MAX_CAND; ++k) accS[k] = accR[k] = 0.0;
      if (ex < n - 1)
        for (int j = 0; j < run; ++j)
          if (ey0 + j < n - 1)
            energy_elem<float>(Rp, Tp, px, py, zx, zy, ox_, oy_, n, h, width, st[pair], ex, ey0 + j, accS, accR);
      energy_write_partials(accS, accR, st[pair].ncand, part, pair, tile, tiles);
    }
  }
}

// ================================================================================================
// SSD variation DE = F_SSD + lambda K phi (k_grad<S, DT_SSD> in FP32 arithmetic).
// ================================================================================================
__global__ void __launch_bounds__(256)
k_grad_f32(const float *__restrict__ R, const float *__restrict__ T, const float *__restrict__ phi,
           float *__restrict__ DE, int n, double vol_d, const PairState *st, double lambda_default, int ntx, int nty,
           int npairs) {
#pragma clang fp contract(fast)
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  if (st && !participates(st[pair])) return;
  const double lambda6 = (st ? st[pair].lambda : lambda_default) * (1.0 / 6.0);

  __shared__ float cf[GR_EY * GR_EX][4][2];  // force contribution per element, dof, component
  __shared__ float ck[GR_EY * GR_EX][4][2];  // stiffness contribution

  const size_t nn = (size_t)n * n;
  const float *Rp = R + pair * nn, *Tp = T + pair * nn;
  const __amdgpu_buffer_rsrc_t Trs = image_rsrc(Tp, (int)nn);
  const float *px = phi + pair * 2 * nn, *py = px + nn;
  const int x0 = (tile % ntx) * GR_TX, y0 = (tile / ntx) * GR_TY;
  const float nm1 = (float)(n - 1);
  const float wv = (float)(0.25 * vol_d);  // quad weight * vol

  for (int e = threadIdx.x; e < GR_EX * GR_EY; e += 256) {
    const int ex = x0 - 1 + (e % GR_EX), ey = y0 - 1 + (e / GR_EX);
    if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
    float f0[4], f1[4], rn[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int g = l2g(ex, ey, b, n);
      f0[b] = px[g];
      f1[b] = py[g];
      rn[b] = Rp[g];
    }
    float nl[4][2], ox4[4], oy4[4], r4[4];
    quad_values(f0, ox4);
    quad_values(f1, oy4);
    quad_values(rn, r4);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const HitF hp = hit_rel(ex, ey, ox4[q] * nm1 + qpxf(q), oy4[q] * nm1 + qpyf(q), n);
      float t4[4];
      const float diff = bilin_f(Trs, hp, n, t4) - r4[q];
      // gradient of the bilinear T at the hit, in domain units
      const float ax = t4[1] - t4[0], bx = t4[3] - t4[2];
      const float ay = t4[2] - t4[0], by = t4[3] - t4[1];
      const float gx = (hp.ly * (bx - ax) + ax) * nm1;
      const float gy = (hp.lx * (by - ay) + ay) * nm1;
      nl[q][0] = hp.inx ? gx * diff : 0.f;
      nl[q][1] = hp.iny ? gy * diff : 0.f;
    }
#pragma unroll
    for (int dof = 0; dof < 4; ++dof) {
      float a0 = 0.f, a1 = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a0 += bqf(dof, q) * nl[q][0];
        a1 += bqf(dof, q) * nl[q][1];
      }
      cf[e][dof][0] = wv * a0;
      cf[e][dof][1] = wv * a1;
    }
    // 6 K_loc = [4 -1 -1 -2; -1 4 -2 -1; -1 -2 4 -1; -2 -1 -1 4] applied to the element's dofs. Its
    // rows sum to zero, so row i is sum over j != i of K_ij (f_j - f_i): differences of close values
    // instead of a sum that cancels to the FP32 ulp of f for smooth fields. The integer weights are
    // exact; the factor lambda / 6 is applied in FP64 (FP32's 1/6 carries a 3e-8 bias).
    const float kl[4][4] = {{4.f, -1.f, -1.f, -2.f}, {-1.f, 4.f, -2.f, -1.f}, {-1.f, -2.f, 4.f, -1.f},
                            {-2.f, -1.f, -1.f, 4.f}};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float s0 = 0.f, s1 = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j == i) continue;
        s0 += kl[i][j] * (f0[j] - f0[i]);
        s1 += kl[i][j] * (f1[j] - f1[i]);
      }
      ck[e][i][0] = s0;
      ck[e][i][1] = s1;
    }
  }
  __syncthreads();
  const int x = x0 + (threadIdx.x % GR_TX), y = y0 + (threadIdx.x / GR_TX);
  if (x >= n || y >= n) return;
  float F0 = 0.f, F1 = 0.f, K0 = 0.f, K1 = 0.f;
#pragma unroll
  for (int ey = y - 1; ey <= y; ++ey)
#pragma unroll
    for (int ex = x - 1; ex <= x; ++ex) {
      if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
      const int e = (ex - (x0 - 1)) + (ey - (y0 - 1)) * GR_EX;
      const int dof = (x - ex) + 2 * (y - ey);
      F0 += cf[e][dof][0];
      F1 += cf[e][dof][1];
      K0 += ck[e][dof][0];
      K1 += ck[e][dof][1];
    }
  float *o = DE + pair * 2 * nn;
  const int i = x + y * n;
  o[i] = (float)(lambda6 * (double)K0 + (double)F0);
  o[nn + i] = (float)(lambda6 * (double)K1 + (double)F1);
}

// ================================================================================================
// V-cycle of the H1 metric in FP32 arithmetic (k_mg_down / k_mg_up / k_mg_coarse).
// ================================================================================================
__device__ __forceinline__ void level_consts_f(const PairState *st, int pair, double sigma_default, int n, float &hs,
                                               float &inv_diag) {
  const double sigma = st ? st[pair].sigma : sigma_default;
  const double tau = 0.5 * (sigma * sigma);
  const double s = (double)((n - 1) * (n - 1));
  hs = (float)(tau * s);
  inv_diag = (float)(1.0 / (1 + 4 * tau * s));
}

// A u of the FD heat operator (heat_apply) in FP32. The stencil 4u - (neighbours) is summed as
// differences u - u_nb: hs grows like n^2, and 4u - sum(u_nb) of a smooth field would cancel to a
// few FP32 ulps of u (a 1e-4 relative error of A u at n = 256), while each difference of close
// values is exact.
template <typename Get>
__device__ __forceinline__ float heat_apply_f(const Get &get, int x, int y, int n, float hs) {
#pragma clang fp contract(fast)
  if (n == 1) return get(0, 0);
  const float u = get(x, y);
  const int e = n - 1;
  // Neighbours outside the grid are mirrored inward, so on an edge the inward difference appears
  // twice (u - v) + (u - v) = 2 (u - v) exactly, and every node takes the same branch-free form:
  // interior ((u - r) + (u - l)) + ((u - t) + (u - b)); edge, e.g. y = 0, 2 (u - t) + ((u - r) + (u - l))
  // summed in the other order (FP addition commutes); corner 2 (u - r) + 2 (u - t) = 2 ((u - r) +
  // (u - t)) exactly. A wave covering a grid edge then runs one path instead of three.
  const int xl = x > 0 ? x - 1 : x + 1, xr = x < e ? x + 1 : x - 1;
  const int yb = y > 0 ? y - 1 : y + 1, yt = y < e ? y + 1 : y - 1;
  const float lap = ((u - get(xr, y)) + (u - get(xl, y))) + ((u - get(x, yt)) + (u - get(x, yb)));
  return hs * lap + u;
}

template <int TX>
__global__ void __launch_bounds__(256)
k_mg_down_f32(const float *__restrict__ rhs, int negate, float *__restrict__ x2out, float *__restrict__ rhsc, int n,
              const PairState *st, double sigma_default, int ntx, int nty, int nfields) {
#pragma clang fp contract(fast)
  constexpr int HX = TX + 4;
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * nfields);
  const int field = wg / tiles, tile = wg - field * tiles;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  float hs, idg;
  level_consts_f(st, pair, sigma_default, n, hs, idg);

  __shared__ float sr[MG32_HY][HX];    // rhs on tile+2
  __shared__ float s1[MG32_HY][HX];    // x1 on tile+2
  __shared__ float s2[MG32_HY][HX];    // x2 on tile+1
  __shared__ float sres[MG32_TY][TX];  // residual on tile
  const size_t nn = (size_t)n * n;
  const float *rp = rhs + field * nn;
  const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * MG32_TY;
  const int bx = x0 - 2, by = y0 - 2;
  const float sg = negate ? -1.f : 1.f;

  // every load of the halo tile issued before the first is used (clamped indices; cells outside
  // the grid are not stored), so a workgroup waits for one memory round trip, not one per sweep
  constexpr int NH = HX * MG32_HY, PER = (NH + 255) / 256;
  float hv[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = threadIdx.x + 256 * r;
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    const bool ok = i < NH && gx >= 0 && gy >= 0 && gx < n && gy < n;
    hv[r] = rp[ok ? gx + gy * n : 0];
  }
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = threadIdx.x + 256 * r;
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    if (i < NH && gx >= 0 && gy >= 0 && gx < n && gy < n) {
      const float v = sg * hv[r];
      sr[ly][lx] = v;
      s1[ly][lx] = v * idg;  // first Jacobi sweep from zero
    }
  }
  __syncthreads();
  auto get1 = [&](int gx, int gy) { return s1[gy - by][gx - bx]; };
  for (int i = threadIdx.x; i < HX * MG32_HY; i += 256) {
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    if (lx >= 1 && ly >= 1 && lx < HX - 1 && ly < MG32_HY - 1 && gx >= 0 && gy >= 0 && gx < n && gy < n)
      s2[ly][lx] = (sr[ly][lx] - heat_apply_f(get1, gx, gy, n, hs)) * idg + s1[ly][lx];
  }
  __syncthreads();
  auto get2 = [&](int gx, int gy) { return s2[gy - by][gx - bx]; };
  float *xo = x2out + field * nn;
  for (int i = threadIdx.x; i < TX * MG32_TY; i += 256) {
    const int tx = i % TX, ty = i / TX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx < n && gy < n) {
      sres[ty][tx] = sr[ty + 2][tx + 2] - heat_apply_f(get2, gx, gy, n, hs);
      xo[gx + gy * n] = s2[ty + 2][tx + 2];
    }
  }
  __syncthreads();
  const int nc = n / 2;
  float *rc = rhsc + field * (size_t)nc * nc;
  for (int i = threadIdx.x; i < (TX / 2) * (MG32_TY / 2); i += 256) {
    const int cx = i % (TX / 2), cy = i / (TX / 2);
    const int gcx = x0 / 2 + cx, gcy = y0 / 2 + cy;
    if (gcx < nc && gcy < nc)
      rc[gcx + gcy * nc] =
          0.25f * ((sres[2 * cy][2 * cx] + sres[2 * cy][2 * cx + 1]) + (sres[2 * cy + 1][2 * cx] + sres[2 * cy + 1][2 * cx + 1]));
  }
}

template <int TX>
__global__ void __launch_bounds__(256)
k_mg_up_f32(const float *__restrict__ x2in, const float *__restrict__ xc, const float *__restrict__ rhs, int negate,
            float *__restrict__ xout, int n, const PairState *st, double sigma_default, int ntx, int nty, int nfields,
            const float *__restrict__ dold, double *__restrict__ nrm_part) {
#pragma clang fp contract(fast)
  constexpr int HX = TX + 4;
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * nfields);
  const int field = wg / tiles, tile = wg - field * tiles;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  float hs, idg;
  level_consts_f(st, pair, sigma_default, n, hs, idg);

  __shared__ float sx[MG32_HY][HX];  // x on tile+2
  __shared__ float sr[MG32_HY][HX];  // rhs on tile+1
  __shared__ float s3[MG32_HY][HX];  // x3 on tile+1
  __shared__ double red[4 * NORM_TERMS];
  const size_t nn = (size_t)n * n;
  const int nc = n / 2;
  const float *x2p = x2in + field * nn, *rp = rhs + field * nn, *xcp = xc + field * (size_t)nc * nc;
  const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * MG32_TY;
  const int bx = x0 - 2, by = y0 - 2;
  const float sg = negate ? -1.f : 1.f;
  // the previous direction at this thread's output nodes, loaded now so the load overlaps the sweeps
  constexpr int OUT_PER_THREAD = TX * MG32_TY / 256;
  float dprev[OUT_PER_THREAD];
  if (nrm_part) {  // uniform: every load issued at once, from clamped indices
#pragma unroll
    for (int k = 0; k < OUT_PER_THREAD; ++k) {
      const int i = threadIdx.x + 256 * k;
      const int gx = x0 + i % TX, gy = y0 + i / TX;
      const float v = dold[field * nn + ((gx < n && gy < n) ? gx + gy * n : 0)];
      dprev[k] = (gx < n && gy < n) ? v : 0.f;
    }
  } else {
#pragma unroll
    for (int k = 0; k < OUT_PER_THREAD; ++k) dprev[k] = 0.f;
  }

  // every load of the halo tile issued before the first is used (clamped indices; cells outside
  // the grid are not stored)
  constexpr int NH = HX * MG32_HY, PER = (NH + 255) / 256;
  float hx2[PER], hxc[PER], hr[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = threadIdx.x + 256 * r;
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    const bool ok = i < NH && gx >= 0 && gy >= 0 && gx < n && gy < n;
    const int gi = ok ? gx + gy * n : 0, ci = ok ? (gx >> 1) + (gy >> 1) * nc : 0;
    hx2[r] = x2p[gi];
    hxc[r] = xcp[ci];
    hr[r] = rp[gi];
  }
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = threadIdx.x + 256 * r;
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    if (i < NH && gx >= 0 && gy >= 0 && gx < n && gy < n) {
      sx[ly][lx] = hx2[r] + hxc[r];
      sr[ly][lx] = sg * hr[r];
    }
  }
  __syncthreads();
  auto getx = [&](int gx, int gy) { return sx[gy - by][gx - bx]; };
  for (int i = threadIdx.x; i < HX * MG32_HY; i += 256) {
    const int lx = i % HX, ly = i / HX;
    const int gx = bx + lx, gy = by + ly;
    if (lx >= 1 && ly >= 1 && lx < HX - 1 && ly < MG32_HY - 1 && gx >= 0 && gy >= 0 && gx < n && gy < n)
      s3[ly][lx] = (sr[ly][lx] - heat_apply_f(getx, gx, gy, n, hs)) * idg + sx[ly][lx];
  }
  __syncthreads();
  auto get3 = [&](int gx, int gy) { return s3[gy - by][gx - bx]; };
  float *xo = xout + field * nn;
  double acc[NORM_TERMS] = {0., 0., 0., 0.};  // <z,z>, <DE,z>, <DE,d_old>, <DE,DE> (DE = the raw input)
#pragma unroll
  for (int k = 0; k < OUT_PER_THREAD; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int tx = i % TX, ty = i / TX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx < n && gy < n) {
      const float o = (sr[ty + 2][tx + 2] - heat_apply_f(get3, gx, gy, n, hs)) * idg + s3[ty + 2][tx + 2];
      xo[gx + gy * n] = o;
      if (nrm_part) {
        const double od = (double)o, de = (double)(sg * sr[ty + 2][tx + 2]);  // the raw input
        acc[0] += od * od;
        acc[1] += de * od;
        acc[2] += de * (double)dprev[k];
        acc[3] += de * de;
      }
    }
  }
  if (nrm_part) {
    block_sum_n<NORM_TERMS>(acc, red);
    if (threadIdx.x == 0)
      for (int t = 0; t < NORM_TERMS; ++t) nrm_part[((size_t)field * NORM_TERMS + t) * tiles + tile] = acc[t];
  }
}

// Whole V-cycle below (and including) a level of at most 64x64 nodes in one workgroup of
// MGC_THREADS threads (k_mg_coarse, one level deeper: FP32 arrays of level 6 and below fit in LDS).
// The Jacobi sweeps write a second buffer instead of updating in place, the residual is restricted as
// it is formed and the prolongation is read on the fly by the first sweep of the up leg, so a level
// costs 3 barriers down and 2 up. Node coordinates come from shifts and masks (a division by the
// runtime level size would cost an integer-division sequence per node).
constexpr int MGC_THREADS = 1024;
constexpr int MGC_MAXL_F32 = 6;

__global__ void __launch_bounds__(MGC_THREADS)
k_mg_coarse_f32(const float *__restrict__ rhs, int negate, float *__restrict__ xout, int L, const PairState *st,
                double sigma_default, int nfields, const float *__restrict__ dold, double *__restrict__ nrm_part) {
#pragma clang fp contract(fast)
  const int field = blockIdx.x;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  const double sigma = st ? st[pair].sigma : sigma_default;
  const double tau = 0.5 * (sigma * sigma);
  __shared__ float X[MGC_OFF[MGC_MAXL_F32 + 1]], B[MGC_OFF[MGC_MAXL_F32 + 1]], Y[1 << (2 * MGC_MAXL_F32)];
  __shared__ double red[(MGC_THREADS / 64) * NORM_TERMS];
  const int n = 1 << L;
  const size_t nn = (size_t)n * n;
  const float *rp = rhs + field * nn;
  const float sg = negate ? -1.f : 1.f;
  constexpr int PER = (1 << (2 * MGC_MAXL_F32)) / MGC_THREADS;  // nodes per thread at the largest level
  {
    float v[PER];  // every load issued before the first LDS store (indices clamped into the field)
#pragma unroll
    for (int k = 0; k < PER; ++k) v[k] = rp[min((int)threadIdx.x + MGC_THREADS * k, (int)nn - 1)];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = threadIdx.x + MGC_THREADS * k;
      if (i < (int)nn) B[MGC_OFF[L] + i] = sg * v[k];
    }
  }
  __syncthreads();

  auto consts = [&](int l, float &hs, float &idg) {
    const int m = 1 << l;
    const double s = (double)((m - 1) * (m - 1));
    hs = (float)(tau * s);
    idg = (float)(1.0 / (1 + 4 * tau * s));
  };

  // Down leg of level l with nthr threads (thread tid): two Jacobi sweeps from zero, the second kept
  // in X for the up leg, then the residual restricted to level l - 1.
  auto down_level = [&](int l, int tid, int nthr, auto sync) {
    const int m = 1 << l, mm = m * m, mc = m / 2;
    float hs, idg;
    consts(l, hs, idg);
    float *x = X + MGC_OFF[l];
    const float *b = B + MGC_OFF[l];
    for (int i = tid; i < mm; i += nthr) Y[i] = b[i] * idg;  // first sweep from zero, into Y
    sync();
    auto gy = [&](int gx, int gyy) { return Y[gx + gyy * m]; };
    for (int i = tid; i < mm; i += nthr) x[i] = (b[i] - heat_apply_f(gy, i & (m - 1), i >> l, m, hs)) * idg + Y[i];
    sync();
    auto gx2 = [&](int px, int py) { return x[px + py * m]; };
    for (int i = tid; i < mc * mc; i += nthr) {  // residual at the 4 children of a coarse node, averaged
      const int cx = i & (mc - 1), cy = i >> (l - 1);
      float r[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int fx = 2 * cx + (c & 1), fy = 2 * cy + (c >> 1);
        r[c] = b[fx + fy * m] - heat_apply_f(gx2, fx, fy, m, hs);
      }
      B[MGC_OFF[l - 1] + i] = 0.25f * ((r[0] + r[1]) + (r[2] + r[3]));
    }
    sync();
  };
  // Up leg of level l: the injected coarse correction read on the fly by the first sweep, the
  // second sweep into X.
  auto up_level = [&](int l, int tid, int nthr, auto sync) {
    const int m = 1 << l, mm = m * m, mc = m / 2;
    float hs, idg;
    consts(l, hs, idg);
    float *x = X + MGC_OFF[l];
    const float *xc = X + MGC_OFF[l - 1];
    const float *b = B + MGC_OFF[l];
    auto gp = [&](int px, int py) { return x[px + py * m] + xc[(px >> 1) + (py >> 1) * mc]; };
    for (int i = tid; i < mm; i += nthr) {
      const int px = i & (m - 1), py = i >> l;
      Y[i] = (b[i] - heat_apply_f(gp, px, py, m, hs)) * idg + gp(px, py);
    }
    sync();
    auto gy = [&](int gx, int gyy) { return Y[gx + gyy * m]; };
    for (int i = tid; i < mm; i += nthr) x[i] = (b[i] - heat_apply_f(gy, i & (m - 1), i >> l, m, hs)) * idg + Y[i];
    sync();
  };
  auto block_sync = [] { __syncthreads(); };
  // Levels of at most 64 nodes run in wave 0 alone: its instructions, LDS accesses included, execute
  // in order for all lanes, so a wavefront fence (a compiler barrier) replaces the workgroup barrier.
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  constexpr int WAVE_LEVELS = 3;  // 8x8 nodes and below
  for (int l = L; l > WAVE_LEVELS; --l) down_level(l, (int)threadIdx.x, MGC_THREADS, block_sync);
  if (threadIdx.x < 64) {
    const int lw = L < WAVE_LEVELS ? L : WAVE_LEVELS;
    for (int l = lw; l >= 1; --l) down_level(l, (int)threadIdx.x, 64, wave_sync);
    if (threadIdx.x == 0) X[0] = B[0];  // 1x1 level: A = 1, the sweeps converge in one step
    wave_sync();
    for (int l = 1; l <= lw; ++l) up_level(l, (int)threadIdx.x, 64, wave_sync);
  }
  __syncthreads();
  for (int l = WAVE_LEVELS + 1; l <= L; ++l) up_level(l, (int)threadIdx.x, MGC_THREADS, block_sync);
  float *xo = xout + field * nn;
  double acc[NORM_TERMS] = {0., 0., 0., 0.};  // <z,z>, <DE,z>, <DE,d_old>, <DE,DE> (DE = the raw input)
  float rv[PER], ov[PER];
  if (nrm_part) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = min((int)threadIdx.x + MGC_THREADS * k, (int)nn - 1);
      rv[k] = rp[i];
      ov[k] = dold[field * nn + i];
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + MGC_THREADS * k;
    if (i < (int)nn) {
      const float o = X[MGC_OFF[L] + i];
      xo[i] = o;
      if (nrm_part) {
        const double od = (double)o, de = (double)rv[k];
        acc[0] += od * od;
        acc[1] += de * od;
        acc[2] += de * (double)ov[k];
        acc[3] += de * de;
      }
    }
  }
  if (nrm_part) {
    block_sum_n<NORM_TERMS, MGC_THREADS>(acc, red);
    if (threadIdx.x == 0)
      for (int t = 0; t < NORM_TERMS; ++t) nrm_part[(size_t)field * NORM_TERMS + t] = acc[t];
  }
}

}  // namespace msg
