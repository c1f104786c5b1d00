// Phase correlation of two frames on the GPU (PhaseCorrelationRegistration::registerImages,
// quoc/paramReg.h:2057-2096, over Convolution::phaseCorrelation, quoc/convolution.h:1049-1114):
// a 2-D FP64 FFT built from one-workgroup row transforms (radix-2 in LDS) and tiled transposes,
// the normalised cross power spectrum, the inverse transform and the peak search.
#pragma once

#include <hip/hip_runtime.h>

namespace msg {

constexpr int FFT_MAXLOG = 12;  // a row of up to 4096 complex doubles fits the 64 KiB of LDS

// w[k] = exp(-2 pi i k / n), k < n/2
__global__ void k_fft_twiddles(double2 *w, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n / 2) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    w[k] = make_double2(c, s);
  }
}

// One row of n = 2^logn points per workgroup: bit-reversed load into LDS, logn radix-2 butterfly
// stages in place, store. Unnormalised in both directions (FFTW's convention); the inverse
// conjugates the twiddles. In is a real scalar type (imaginary part 0) or double2.
template <typename In>
__global__ __launch_bounds__(1024) void k_fft_rows(const In *__restrict__ in, double2 *__restrict__ out,
                                                   const double2 *__restrict__ w, int n, int logn, int inverse) {
  extern __shared__ double2 buf[];
  const size_t base = (size_t)blockIdx.x * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int r = (int)(__brev((unsigned)i) >> (32 - logn));
    if constexpr (sizeof(In) == sizeof(double2)) {
      buf[r] = in[base + i];
    } else {
      buf[r] = make_double2((double)in[base + i], 0.0);
    }
  }
  __syncthreads();
  for (int s = 1; s <= logn; ++s) {
    const int half = 1 << (s - 1), tstride = n >> s;
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const int j = b & (half - 1);
      const int i0 = ((b >> (s - 1)) << s) + j, i1 = i0 + half;
      double2 t = w[j * tstride];
      if (inverse) t.y = -t.y;
      const double2 a = buf[i0], c = buf[i1];
      const double2 ct = make_double2(c.x * t.x - c.y * t.y, c.x * t.y + c.y * t.x);
      buf[i0] = make_double2(a.x + ct.x, a.y + ct.y);
      buf[i1] = make_double2(a.x - ct.x, a.y - ct.y);
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[base + i] = buf[i];
}

// out = in^T for an n x n complex array; 32 x 32 tiles through LDS, 32 x 8 threads.
__global__ __launch_bounds__(256) void k_transpose_c(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                     int n) {
  __shared__ double2 tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32, tx = threadIdx.x, ty = threadIdx.y;
  for (int k = ty; k < 32; k += 8) {
    const int x = bx + tx, y = by + k;
    if (x < n && y < n) tile[k][tx] = in[(size_t)y * n + x];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int x = by + tx, y = bx + k;
    if (x < n && y < n) out[(size_t)y * n + x] = tile[tx][k];
  }
}

// a <- a * conj(b), scaled to unit modulus when `normalise` (moduli below 1e-15 give 0), in the
// reference's operation order (convolution.h:1095-1102).
__global__ void k_cross_power(double2 *__restrict__ a, const double2 *__restrict__ b, size_t cnt, int normalise) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    const double2 g2 = a[i], g1 = b[i];
    const double tx = g2.x * g1.x + g2.y * g1.y, ty = g2.y * g1.x - g2.x * g1.y;
    double scale = 1.0;
    if (normalise) {
      const double nrm = sqrt(tx * tx + ty * ty);
      scale = fabs(nrm) < 1e-15 ? 0.0 : 1.0 / nrm;
    }
    a[i] = make_double2(tx * scale, ty * scale);
  }
}

// out = Re(in) * f
__global__ void k_real_part(const double2 *__restrict__ in, double *__restrict__ out, size_t cnt, double f) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i].x * f;
}

// Per workgroup of 256: the largest value and its first (lowest) index.
__global__ __launch_bounds__(256) void k_argmax_part(const double *__restrict__ v, size_t cnt, double *__restrict__ bv,
                                                     long long *__restrict__ bi) {
  __shared__ double sv[256];
  __shared__ long long si[256];
  double m = -INFINITY;
  long long mi = -1;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    const double x = v[i];
    if (mi < 0 || x > m) {  // indices grow along the loop: the first maximum is kept
      m = x;
      mi = (long long)i;
    }
  }
  sv[threadIdx.x] = m;
  si[threadIdx.x] = mi;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double x = sv[threadIdx.x + o];
      const long long xi = si[threadIdx.x + o];
      if (xi >= 0 && (si[threadIdx.x] < 0 || x > sv[threadIdx.x] || (x == sv[threadIdx.x] && xi < si[threadIdx.x]))) {
        sv[threadIdx.x] = x;
        si[threadIdx.x] = xi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bv[blockIdx.x] = sv[0];
    bi[blockIdx.x] = si[0];
  }
}

// The (2m+1)^2 values around the origin, periodically wrapped, x fastest from -m (m < n).
__global__ void k_gather_window(const double *__restrict__ v, int n, int m, double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, side = 2 * m + 1;
  if (k >= side * side) return;
  const int dx = k % side - m, dy = k / side - m;
  out[k] = v[(size_t)((dy + n) % n) * n + (dx + n) % n];
}

// A displacement field constant (tx, ty) over n*n nodes (x component, then y)
template <typename S>
__global__ void k_fill_translation(S *__restrict__ phi, size_t nn, double tx, double ty) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
    phi[i] = (S)tx;
    phi[nn + i] = (S)ty;
  }
}

}  // namespace msg
