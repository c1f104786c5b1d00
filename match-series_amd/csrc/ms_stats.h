// Reconstruction statistics and input normalisation kernels.
#pragma once

#include "ms_device.h"

namespace msg {

// Per pixel over nimg warped frames (+Inf = no sample): mean of the finite samples in frame order,
// sample count, and median (SeriesMatching::averageSeries, matchSeries.h:404-444;
// MultiVector::getMedianVecOverComponents, aol/multiVector.cpp:99-117; Vector::getMedianValue,
// aol/vec.cpp:933-949: odd count -> middle element, even -> mean of the two middle elements).
// One wave64 per P pixels, 64/P lanes per pixel; the samples of the P pixels live in LDS (P * nimg
// doubles, at most 64 KiB) and the order statistics are found by counting, which needs no sort and
// no per-lane arrays: lane l of a pixel tests the candidates i = l, l + 64/P, ... and the lanes agree
// by shuffles (the k-th smallest value is unique, whichever candidate carries it).
constexpr int ST_LDS_DOUBLES = 8192;           // 64 KiB of samples per wave
constexpr int ST_MAXF = ST_LDS_DOUBLES;         // P = 1: one pixel per wave

// Pixels per wave for a series of nimg frames (64, 16, 4 or 1), or 0 beyond ST_MAXF.
inline int stats_pixels_per_wave(int nimg) {
  for (int p = 64; p >= 1; p /= 4)
    if ((long long)p * nimg <= ST_LDS_DOUBLES) return p;
  return 0;
}

template <typename S, int P>
__global__ void __launch_bounds__(64)
k_series_stats(const S *__restrict__ stack, int nimg, int nn, double fill, S *__restrict__ mean,
               S *__restrict__ count, S *__restrict__ median) {
  constexpr int L = 64 / P;  // lanes per pixel
  extern __shared__ double sv[];  // [nimg][P]
  const int q = threadIdx.x / L, l = threadIdx.x % L;
  const int j = blockIdx.x * P + q;
  const bool valid = j < nn;
  for (int f = l; f < nimg; f += L)
    sv[f * P + q] = valid ? ld(stack + (size_t)f * nn, j) : __builtin_inf();
  __syncthreads();
  if (!valid) return;
  double m = 0;
  int c = 0;
  for (int f = 0; f < nimg; ++f) {  // every lane of the pixel sums in frame order
    const double v = sv[f * P + q];
    if (v != __builtin_inf()) {
      m += v;
      ++c;
    }
  }
  m = (c > 0) ? m / c : fill;
  double med = fill;
  if (c > 0) {
    // k-th smallest finite sample: the value v_i with #(v < v_i) <= k < #(v <= v_i)
    const int k2 = c / 2, k1 = (c % 2 == 1) ? k2 : k2 - 1;
    double a = 0, b = 0;
    bool fa = false, fb = false;
    for (int i = l; i < nimg && !(fa && fb); i += L) {
      const double vi = sv[i * P + q];
      if (vi == __builtin_inf()) continue;
      int lt = 0, le = 0;
      for (int f = 0; f < nimg; ++f) {
        const double v = sv[f * P + q];
        if (v == __builtin_inf()) continue;
        lt += (v < vi);
        le += (v <= vi);
      }
      if (!fa && lt <= k1 && k1 < le) {
        a = vi;
        fa = true;
      }
      if (!fb && lt <= k2 && k2 < le) {
        b = vi;
        fb = true;
      }
    }
    // the lanes of one pixel are L consecutive lanes: butterfly over them
    for (int o = L / 2; o > 0; o /= 2) {
      const double oa = __shfl_xor(a, o), ob = __shfl_xor(b, o);
      const int ofa = __shfl_xor((int)fa, o), ofb = __shfl_xor((int)fb, o);
      if (!fa && ofa) {
        a = oa;
        fa = true;
      }
      if (!fb && ofb) {
        b = ob;
        fb = true;
      }
    }
    med = (c % 2 == 1) ? b : (a + b) / 2;
  }
  if (l == 0) {
    mean[j] = static_cast<S>(m);
    count[j] = static_cast<S>((double)c);
    median[j] = static_cast<S>(med);
  }
}

// prepareMultilevelArray (registration.h:331-347) given the sorted copy of the image:
// pct >= 0: clamp to [lo, hi], subtract the minimum (lo) and divide by the maximum (hi - lo);
// pct < 0: subtract the minimum if negative or min_to_zero, divide by the (shifted) maximum.
template <typename S>
__global__ void k_normalize(S *__restrict__ img, int nn, const S *__restrict__ sorted, int i0, int i1, int saturate,
                            int min_to_zero) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nn) return;
  double v = ld(img, j);
  if (saturate) {
    const double lo = ld(sorted, i0), hi = ld(sorted, i1);
    v = v < lo ? lo : (v > hi ? hi : v);
    v += -1 * lo;                 // addToAll(-1 * getMinValue()), the minimum being lo
    const double mx = hi + -1 * lo;  // getMaxValue() of the shifted image
    if (mx != 0) v /= mx;
  } else {
    const double mn = ld(sorted, 0), mxo = ld(sorted, nn - 1);
    double mx = mxo;
    if (mn < 0. || min_to_zero) {
      v += -mn;
      mx = mxo + -mn;
    }
    if (mx != 0.) v /= mx;
  }
  img[j] = static_cast<S>(v);
}

}  // namespace msg
