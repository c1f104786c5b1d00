// Host side of libmsgpu.so: contexts, launches and the C ABI declared in include/ms_gpu.h.
#include "ms_gpu.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ms_kernels.h"
#include "ms_fast.h"
#include "ms_compose.h"
#include "ms_stats.h"
#include "ms_fft.h"
#include "ms_inverse.h"
#include "ms_xshift.h"

using msg::PairState;
namespace {
thread_local std::string g_err;
}
namespace msg {
// Records an error of another translation unit of the library for msg_last_error().
int set_error(int code, const std::string &m) {
  g_err = m;
  return code;
}
}  // namespace msg

namespace {

int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

#define HIPCHK(x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) return fail(MSG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)                 \
  do {                         \
    int r_ = (x);              \
    if (r_ != MSG_OK) return r_; \
  } while (0)

constexpr int MAXL = 15;

inline int cdiv(int a, int b) { return (a + b - 1) / b; }
inline size_t esz(int dtype) { return dtype == MSG_F64 ? 8 : 4; }
inline hipStream_t strm(void *s) { return reinterpret_cast<hipStream_t>(s); }

int check_dtype(int dtype) {
  if (dtype != MSG_F32 && dtype != MSG_F64) return fail(MSG_ERR_ARG, "dtype must be MSG_F32 or MSG_F64");
  return MSG_OK;
}
int check_n(int n) {
  if (n < 2 || (n & (n - 1)) != 0 || n > (1 << MAXL))
    return fail(MSG_ERR_ARG, "grid size must be a power of two between 2 and 32768");
  return MSG_OK;
}
int log2i(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// ------------------------------------------------------------------------------------------------
// Grow-only device scratch for the stateless operator entry points (one set per device).
// ------------------------------------------------------------------------------------------------
// Per host thread and device, so that concurrent callers (one context per thread) never share it.
struct Scratch {
  std::vector<std::pair<void *, size_t>> bufs;  // slot -> (ptr, bytes)
};
thread_local Scratch g_scratch[16];

int scratch_get(int slot, size_t bytes, void **p) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  Scratch &s = g_scratch[dev & 15];
  if ((int)s.bufs.size() <= slot) s.bufs.resize(slot + 1, {nullptr, 0});
  auto &b = s.bufs[slot];
  if (b.second < bytes) {
    if (b.first) HIPCHK(hipFree(b.first));
    b.first = nullptr;
    b.second = 0;
    HIPCHK(hipMalloc(&b.first, bytes));
    b.second = bytes;
  }
  *p = b.first;
  return MSG_OK;
}

// Pinned snapshots of small device flags with an event per snapshot, per thread and device, for host
// checks that run one batch of launches behind the stream (made on first use, kept).
struct FlagSnap {
  int *h = nullptr;  // 2 snapshots of up to 4 ints
  hipEvent_t ev[2] = {nullptr, nullptr};
};
thread_local FlagSnap g_flagsnap[16];

int flag_snap(FlagSnap **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  FlagSnap &f = g_flagsnap[dev & 15];
  if (!f.h) {
    HIPCHK(hipHostMalloc((void **)&f.h, sizeof(int) * 8, hipHostMallocDefault));
    for (auto &e : f.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *out = &f;
  return MSG_OK;
}

// ------------------------------------------------------------------------------------------------
// Kernel timing with HIP events on the stream each kernel is launched on (opt-in, per class mask).
// ------------------------------------------------------------------------------------------------
enum KernelClass { KC_ENERGY = 0, KC_GRAD = 1, KC_VCYCLE = 2, KC_NCG = 3, KC_UPDATE = 4, KC_CONTROL = 5,
                   KC_COMPOSE = 6, KC_COUNT = 7 };
struct ProfRec {
  int cls;
  hipEvent_t a, b;
  double bytes;
};
struct Profiler {
  std::mutex mu;
  unsigned mask = 0;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  double ms[KC_COUNT] = {};
  long long launches[KC_COUNT] = {};
  double bytes[KC_COUNT] = {};
};
Profiler g_prof;
thread_local bool g_capturing = false;  // no event timing inside a stream capture
int g_default_mode = 0;
int g_default_exact = 1;

hipEvent_t prof_event() {
  if (!g_prof.pool.empty()) {
    hipEvent_t e = g_prof.pool.back();
    g_prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// RAII bracket around one launch: records start/stop events when the class is enabled.
struct ProfScope {
  bool on;
  int cls;
  hipStream_t s;
  double bytes;
  hipEvent_t a = nullptr;
  ProfScope(int c, hipStream_t st, double b) : on(((g_prof.mask >> c) & 1u) && !g_capturing), cls(c), s(st), bytes(b) {
    if (on) {
      std::lock_guard<std::mutex> g(g_prof.mu);
      a = prof_event();
      (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (on) {
      std::lock_guard<std::mutex> g(g_prof.mu);
      hipEvent_t b = prof_event();
      (void)hipEventRecord(b, s);
      g_prof.pending.push_back({cls, a, b, bytes});
    }
  }
};

// Adds algorithmic bytes to a class whose launches cannot know them (device-decided work).
void prof_add_bytes(int cls, double bytes) {
  if (!((g_prof.mask >> cls) & 1u)) return;
  std::lock_guard<std::mutex> g(g_prof.mu);
  g_prof.bytes[cls] += bytes;
}

void prof_resolve() {
  for (auto &r : g_prof.pending) {
    (void)hipEventSynchronize(r.b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    g_prof.ms[r.cls] += ms;
    g_prof.launches[r.cls] += 1;
    g_prof.bytes[r.cls] += r.bytes;
    g_prof.pool.push_back(r.a);
    g_prof.pool.push_back(r.b);
  }
  g_prof.pending.clear();
}

// ------------------------------------------------------------------------------------------------
// Small helper kernels
// ------------------------------------------------------------------------------------------------
__global__ void k_state_init(PairState *st, int npairs, double lambda) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  static_assert(sizeof(PairState) % 8 == 0, "PairState is a whole number of doubles");
  double *w = reinterpret_cast<double *>(st + p);  // zero in place (a local copy would live in scratch)
  for (size_t i = 0; i < sizeof(PairState) / 8; ++i) w[i] = 0.0;
  PairState &s = st[p];
  s.ncand = 1;
  s.cand[0] = 0.0;
  s.dirmode = msg::DIR_NONE;
  s.lambda = lambda;
  s.active = 1;
  s.sigma = 1.0;
}

template <typename D, typename S>
__global__ void k_convert(const S *__restrict__ in, D *__restrict__ out, size_t cnt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = static_cast<D>(in[i]);
}

int convert(const void *in, int din, void *out, int dout, size_t cnt, hipStream_t s) {
  if (cnt == 0) return MSG_OK;
  if (din == dout) {
    HIPCHK(hipMemcpyAsync(out, in, cnt * esz(din), hipMemcpyDeviceToDevice, s));
    return MSG_OK;
  }
  const int grid = (int)((cnt + 255) / 256);
  if (din == MSG_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid), dim3(256), 0, s, (const double *)in, (float *)out, cnt);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid), dim3(256), 0, s, (const float *)in, (double *)out, cnt);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

// ------------------------------------------------------------------------------------------------
// Launch wrappers shared by the context path and the stateless entry points
// ------------------------------------------------------------------------------------------------
struct EnergyGeom {
  int ntx, nty, nblk;
};
// Tiles of an energy pass at grid size n: 64 x 4 elements, or 64 x 4 run for the FP32 SSD pass
// (k_energy_f32 and its FP64 re-evaluation; the run is en32_run(n)).
EnergyGeom energy_geom(int n, bool f32_ssd = false) {
  EnergyGeom g;
  g.ntx = cdiv(n - 1, msg::EN_TX);
  g.nty = cdiv(n - 1, f32_ssd ? 4 * msg::en32_run(n) : msg::EN_TY);
  g.nblk = g.ntx * g.nty;
  return g;
}
template <typename S>
bool is_f32_ssd(int data_term) {
  return std::is_same<S, float>::value && data_term == msg::DT_SSD;
}

// doubles of the energy pass's partial sums for npairs pairs at grid size n (either data term)
size_t energy_part_size(int n, int npairs) {
  // the x-shift pass writes one partial per image row
  return (size_t)npairs * (msg::MAX_CAND * 4 + 2) * std::max(energy_geom(n).nblk, n);
}
// partial sums per (pair, candidate, term) of the energy pass of a data term at grid size n
template <typename S>
int energy_blocks(int n, int data_term) {
  return data_term == msg::DT_XSHIFT ? n : energy_geom(n, is_f32_ssd<S>(data_term)).nblk;
}

// The images of the data term: R, T (SSD, NCC) or, for DT_CONS, the npsi deformations in T.
template <typename S>
int launch_energy(const S *R, const S *T, const S *phi, const S *z, const S *dold, int n, const PairState *st,
                  double *part, int npairs, hipStream_t s, int data_term, double bytes = 0, int npsi = 0,
                  bool exact_pass = true, double yreg = 0) {
  ProfScope ps(KC_ENERGY, s, bytes);
  const EnergyGeom g = energy_geom(n, is_f32_ssd<S>(data_term));
  const double h = 1.0 / (double)(n - 1);
  const int width = (int)((n - 1) * h);
  if (data_term == msg::DT_XSHIFT)
    hipLaunchKernelGGL((msg::k_energy_xshift<S>), dim3(n * npairs), dim3(msg::XS_THREADS), 3 * n * sizeof(double), s,
                       R, T, phi, z, dold, n, h, yreg, st, part, npairs);
  else if (data_term == msg::DT_CONS)
    hipLaunchKernelGGL((msg::k_energy_cons<S>), dim3(g.nblk * npairs), dim3(256), 0, s, T, npsi, phi, z, dold, n, h,
                       width, st, part, g.ntx, g.nty, npairs);
  else if (data_term == msg::DT_NCC)
    hipLaunchKernelGGL((msg::k_energy_ncc<S>), dim3(g.nblk * npairs), dim3(256), 0, s, R, T, phi, z, dold, n, h,
                       st, part, g.ntx, g.nty, npairs);
  else if constexpr (std::is_same<S, float>::value) {
    if (n > 16384) return fail(MSG_ERR_ARG, "FP32 fields: the SSD energy pass addresses grids up to 16384^2");
    if (n <= 4096 && msg::en32_run(n) == 2)
      hipLaunchKernelGGL((msg::k_energy_f32<true, 2>), dim3(g.nblk * npairs), dim3(256), 0, s, R, T, phi, z, dold, n,
                         h * h, st, part, g.ntx, g.nty, npairs);
    else if (n <= 4096)
      hipLaunchKernelGGL((msg::k_energy_f32<true, 4>), dim3(g.nblk * npairs), dim3(256), 0, s, R, T, phi, z, dold, n,
                         h * h, st, part, g.ntx, g.nty, npairs);
    else
      hipLaunchKernelGGL((msg::k_energy_f32<false, 4>), dim3(g.nblk * npairs), dim3(256), 0, s, R, T, phi, z, dold,
                         n, h * h, st, part, g.ntx, g.nty, npairs);
    if (exact_pass) {  // the pairs flagged for an FP64-arithmetic evaluation (none: a check per pair)
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(msg::k_energy_exact_f32, dim3(std::min(g.nblk, 1024)), dim3(256), 0, s, R, T, phi, z, dold,
                         n, h, width, st, part, g.ntx, g.nty, npairs);
    }
  }
  else
    hipLaunchKernelGGL((msg::k_energy<S>), dim3(g.nblk * npairs), dim3(256), 0, s, R, T, phi, z, dold, n, h, width,
                       st, part, g.ntx, g.nty, npairs);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

template <typename S>
int launch_grad(const S *R, const S *T, const S *phi, S *DE, int n, const PairState *st, double lambda, int npairs,
                hipStream_t s, int data_term, int npsi = 0, double yreg = 0) {
  ProfScope ps(KC_GRAD, s, 6.0 * n * n * sizeof(S) * npairs);
  const int ntx = cdiv(n, msg::GR_TX), nty = cdiv(n, msg::GR_TY);
  const double h = 1.0 / (double)(n - 1);
  const int width = (int)((n - 1) * h);
  if (data_term == msg::DT_XSHIFT)
    hipLaunchKernelGGL((msg::k_grad_xshift<S>), dim3(n * npairs), dim3(msg::XS_THREADS), 3 * n * sizeof(double), s, R,
                       T, phi, DE, n, h, yreg, st, lambda, npairs);
  else if (data_term == msg::DT_CONS)
    hipLaunchKernelGGL((msg::k_grad<S, msg::DT_CONS>), dim3(ntx * nty * npairs), dim3(256), 0, s, R, T, phi, DE, n, h,
                       width, st, lambda, ntx, nty, npairs, npsi);
  else if (data_term == msg::DT_NCC)
    hipLaunchKernelGGL((msg::k_grad<S, msg::DT_NCC>), dim3(ntx * nty * npairs), dim3(256), 0, s, R, T, phi, DE, n, h,
                       width, st, lambda, ntx, nty, npairs, 0);
  else if constexpr (std::is_same<S, float>::value)
    hipLaunchKernelGGL(msg::k_grad_f32, dim3(ntx * nty * npairs), dim3(256), 0, s, R, T, phi, DE, n, h * h, st, lambda,
                       ntx, nty, npairs);
  else
    hipLaunchKernelGGL((msg::k_grad<S, msg::DT_SSD>), dim3(ntx * nty * npairs), dim3(256), 0, s, R, T, phi, DE, n, h,
                       width, st, lambda, ntx, nty, npairs, 0);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

// The reference statistics of the NCC energy for every pair of a level (two passes with their
// control phases); the first energy pass of the level follows.
template <typename S>
int ncc_reference_stats(const S *R, int n, PairState *st, double *part, double *part_n, int B,
                        msg::CtrlArgs a, int *flags, msg::IterTrace *tr, hipStream_t s) {
  const EnergyGeom g = energy_geom(n);
  const double h = 1.0 / (double)(n - 1);
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL((msg::k_ncc_ref<S>), dim3(g.nblk * B), dim3(256), 0, s, R, n, h, st, part, g.ntx, g.nty, B, pass);
    HIPCHK(hipGetLastError());
    a.phase = pass ? msg::PH_NCC_REF1 : msg::PH_NCC_REF0;
    a.nblk_energy = g.nblk;
    a.npairs = B;
    hipLaunchKernelGGL(msg::k_control, dim3(B), dim3(msg::CTRL_THREADS), 0, s, st, part, part_n, a, flags, tr);
    HIPCHK(hipGetLastError());
  }
  return MSG_OK;
}

// Buffers of one V-cycle over fields of level l// Buffers of one V-cycle over fields of level l: per multigrid level m < l the coarse rhs, the
// pre-smoothed iterate and the final iterate.
template <typename S>
struct MGBufs {
  S *rhs[MAXL + 1];
  S *x2[MAXL + 1];
  S *x[MAXL + 1];
  S *x2_top;  // pre-smoothed iterate of the top level
};

// Deepest level the single-workgroup coarse V-cycle handles for a field type.
template <typename S>
constexpr int mgc_maxl() {
  return std::is_same<S, float>::value ? msg::MGC_MAXL_F32 : msg::MGC_MAXL;
}

// Partial sums per field and term of the V-cycle's last pass at level l.
// Tiles of the V-cycle's fine passes at grid size n for a field type
template <typename S>
constexpr int mg_tx(int n) {
  return !std::is_same<S, float>::value ? msg::MG_TX : (n >= msg::MG32_WIDE_FROM ? msg::MG32_TX_WIDE : 32);
}
template <typename S>
constexpr int mg_ty() {
  return std::is_same<S, float>::value ? msg::MG32_TY : msg::MG_TY;
}
template <typename S = double>
int mg_norm_blocks(int l, int coarse_max) {
  if (l <= coarse_max) return 1;
  const int n = 1 << l;
  return cdiv(n, mg_tx<S>(n)) * cdiv(n, mg_ty<S>());
}

// The V-cycle kernels of a field type: FP32 arithmetic for FP32 fields, FP64 for FP64 fields.
template <typename S>
void mg_coarse(dim3 g, hipStream_t s, const S *rhs, int negate, S *xout, int L, const PairState *st, double sigma,
               int nfields, const S *dold, double *nrm_part) {
  if constexpr (std::is_same<S, float>::value)
    hipLaunchKernelGGL(msg::k_mg_coarse_f32, g, dim3(msg::MGC_THREADS), 0, s, rhs, negate, xout, L, st, sigma, nfields,
                       dold, nrm_part);
  else
    hipLaunchKernelGGL((msg::k_mg_coarse<S>), g, dim3(256), 0, s, rhs, negate, xout, L, st, sigma, nfields, dold,
                       nrm_part);
}
template <typename S>
void mg_down(dim3 g, hipStream_t s, const S *rhs, int negate, S *x2out, S *rhsc, int n, const PairState *st,
             double sigma, int ntx, int nty, int nfields) {
  if constexpr (std::is_same<S, float>::value) {
    if (mg_tx<S>(n) == msg::MG32_TX_WIDE)
      hipLaunchKernelGGL(msg::k_mg_down_f32<msg::MG32_TX_WIDE>, g, dim3(256), 0, s, rhs, negate, x2out, rhsc, n, st,
                         sigma, ntx, nty, nfields);
    else
      hipLaunchKernelGGL(msg::k_mg_down_f32<32>, g, dim3(256), 0, s, rhs, negate, x2out, rhsc, n, st, sigma, ntx, nty,
                         nfields);
  }
  else
    hipLaunchKernelGGL((msg::k_mg_down<S>), g, dim3(256), 0, s, rhs, negate, x2out, rhsc, n, st, sigma, ntx, nty,
                       nfields);
}
template <typename S>
void mg_up(dim3 g, hipStream_t s, const S *x2in, const S *xc, const S *rhs, int negate, S *xout, int n,
           const PairState *st, double sigma, int ntx, int nty, int nfields, const S *dold, double *nrm_part) {
  if constexpr (std::is_same<S, float>::value) {
    if (mg_tx<S>(n) == msg::MG32_TX_WIDE)
      hipLaunchKernelGGL(msg::k_mg_up_f32<msg::MG32_TX_WIDE>, g, dim3(256), 0, s, x2in, xc, rhs, negate, xout, n, st,
                         sigma, ntx, nty, nfields, dold, nrm_part);
    else
      hipLaunchKernelGGL(msg::k_mg_up_f32<32>, g, dim3(256), 0, s, x2in, xc, rhs, negate, xout, n, st, sigma, ntx,
                         nty, nfields, dold, nrm_part);
  }
  else
    hipLaunchKernelGGL((msg::k_mg_up<S>), g, dim3(256), 0, s, x2in, xc, rhs, negate, xout, n, st, sigma, ntx, nty,
                       nfields, dold, nrm_part);
}

// out = V-cycle(sign * in) for nfields fields of level l (LinearSmoothOp::apply). With nrm_part, the
// last pass also leaves the partial sums <out,out>, <in,out>, <in,dold>, <in,in> per field and block.
template <typename S>
int launch_vcycle(const S *in, int negate, S *out, int l, const MGBufs<S> &b, const PairState *st, double sigma,
                  int nfields, double *nrm_part, const S *dold, hipStream_t s) {
  ProfScope ps(KC_VCYCLE, s, 0);
  constexpr int CM = mgc_maxl<S>();
  if (l <= CM) {
    mg_coarse<S>(dim3(nfields), s, in, negate, out, l, st, sigma, nfields, dold, nrm_part);
    HIPCHK(hipGetLastError());
    return MSG_OK;
  }
  // down leg: levels l .. CM+1
  for (int m = l; m > CM; --m) {
    const int n = 1 << m;
    const int ntx = cdiv(n, mg_tx<S>(n)), nty = cdiv(n, mg_ty<S>());
    const S *src = (m == l) ? in : b.rhs[m];
    const int neg = (m == l) ? negate : 0;
    S *x2 = (m == l) ? b.x2_top : b.x2[m];
    mg_down<S>(dim3(ntx * nty * nfields), s, src, neg, x2, b.rhs[m - 1], n, st, sigma, ntx, nty, nfields);
    HIPCHK(hipGetLastError());
  }
  mg_coarse<S>(dim3(nfields), s, b.rhs[CM], 0, b.x[CM], CM, st, sigma, nfields,
               (const S *)nullptr, (double *)nullptr);
  HIPCHK(hipGetLastError());
  for (int m = CM + 1; m <= l; ++m) {
    const int n = 1 << m;
    const int ntx = cdiv(n, mg_tx<S>(n)), nty = cdiv(n, mg_ty<S>());
    const S *x2 = (m == l) ? b.x2_top : b.x2[m];
    const S *rhs = (m == l) ? in : b.rhs[m];
    const int neg = (m == l) ? negate : 0;
    S *xo = (m == l) ? out : b.x[m];
    mg_up<S>(dim3(ntx * nty * nfields), s, x2, b.x[m - 1], rhs, neg, xo, n, st, sigma, ntx, nty, nfields, dold,
             (m == l) ? nrm_part : (double *)nullptr);
    HIPCHK(hipGetLastError());
  }
  return MSG_OK;
}

template <typename S>
int launch_restrict(const S *fine, S *coarse, int n, int nfields, hipStream_t s) {
  const size_t cnt = (size_t)(n / 2) * (n / 2) * nfields;
  hipLaunchKernelGGL((msg::k_restrict<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, fine, coarse, n,
                     nfields);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

template <typename S>
int launch_prolong(const S *coarse, S *fine, int nc, int nfields, hipStream_t s) {
  const size_t cnt = (size_t)(2 * nc) * (2 * nc) * nfields;
  hipLaunchKernelGGL((msg::k_prolong<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, coarse, fine, nc,
                     nfields);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

// Sampling positions k * (nc-1)/(n-1) of ResampleCopier (scalarArray.h:311-330): the step divided in
// double and held in long double, each product rounded to double.
int resample_coords(int nc, int n, double **dcoord, hipStream_t s) {
  std::vector<double> c(n);
  const long double hx = (long double)((double)(nc - 1) / (double)(n - 1));
  for (int k = 0; k < n; ++k) c[k] = (double)(k * hx);
  void *p = nullptr;
  CHK(scratch_get(15, sizeof(double) * n, &p));
  HIPCHK(hipMemcpyAsync(p, c.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // c is a host temporary
  *dcoord = (double *)p;
  return MSG_OK;
}

template <typename S>
int launch_resample(const S *src, S *dst, int nc, int n, int nfields, const double *coord, hipStream_t s) {
  const size_t cnt = (size_t)n * n * nfields;
  hipLaunchKernelGGL((msg::k_resample<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, src, dst, nc, n,
                     coord, nfields);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

}  // namespace

// =================================================================================================
// Context
// =================================================================================================
struct msg_ctx {
  int L = 0, B = 0, dtype = MSG_F32, device = 0;
  size_t es = 4;
  void *R[MAXL + 1] = {}, *T[MAXL + 1] = {}, *phi[MAXL + 1] = {};
  void *DE = nullptr, *z = nullptr, *d = nullptr, *x2 = nullptr;
  void *mg_rhs[MAXL + 1] = {}, *mg_x2[MAXL + 1] = {}, *mg_x[MAXL + 1] = {};
  double *part_e = nullptr, *part_n = nullptr;
  double *coord = nullptr;
  PairState *st = nullptr;
  int *flags_d = nullptr, *flags_h = nullptr;  // flags_h: two snapshots of B flags (pipelined replays)
  hipEvent_t flag_ev[2] = {nullptr, nullptr};   // recorded after each snapshot's copy
  msg::IterTrace *trace_d = nullptr;
  int trace_cap = 0;
  hipStream_t stream = nullptr;
  long long cnt_iter = 0, cnt_evals = 0, cnt_passes = 0;
  double cnt_node_iter = 0, cnt_node_evals = 0;  // the same weighted by the level's node count
  std::vector<void *> allocs;
  int mode = 0;                 // 0 host-synchronised rounds, 1 device-driven, 2 device-driven + hipGraph
  struct GraphEntry {
    int level, spec, K, R, trace_cap;
    double lambda, stop_eps;
    int max_iter;
    void *trace;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  std::vector<double> level_start;  // per pair: E and lambda*E_reg at the start of the last level solve
  int data_term = msg::DT_SSD;
  int exact_decisions = 1;      // FP32 SSD: near-tie step decisions on FP64-arithmetic values
  double yreg = 1.0;            // DT_XSHIFT: "yRegFactor"
  // reduceDeformations: the deformations (device, npsi fields at level L) the consistency energy reads,
  // and the descent configuration of plain H1GradientDescent
  const void *psi = nullptr;
  int npsi = 0;
  int sigma_fixed = 0, grad_stop = 0;
  double sigma0 = 1.0;
};

namespace {

int ctx_alloc(msg_ctx *c, void **p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes ? bytes : 16));
  c->allocs.push_back(*p);
  return MSG_OK;
}

template <typename S>
MGBufs<S> ctx_mg(msg_ctx *c) {
  MGBufs<S> b;
  for (int m = 0; m <= MAXL; ++m) {
    b.rhs[m] = (S *)c->mg_rhs[m];
    b.x2[m] = (S *)c->mg_x2[m];
    b.x[m] = (S *)c->mg_x[m];
  }
  b.x2_top = (S *)c->x2;
  return b;
}

template <typename S>
int rebuild_pyramid(msg_ctx *c, void **lev, int pair, int nfields_per_pair) {
  for (int l = c->L; l > 0; --l) {
    const int n = 1 << l;
    const size_t fs = (size_t)n * n * nfields_per_pair, cs = (size_t)(n / 2) * (n / 2) * nfields_per_pair;
    if (pair < 0)
      CHK(launch_restrict<S>((const S *)lev[l], (S *)lev[l - 1], n, nfields_per_pair * c->B, c->stream));
    else
      CHK(launch_restrict<S>((const S *)lev[l] + pair * fs, (S *)lev[l - 1] + pair * cs, n, nfields_per_pair,
                             c->stream));
  }
  return MSG_OK;
}

int upload_host(msg_ctx *c, void *dst, const double *src, size_t cnt) {
  if (c->dtype == MSG_F64) {
    HIPCHK(hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(cnt);
    for (size_t i = 0; i < cnt; ++i) tmp[i] = (float)src[i];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), cnt * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return MSG_OK;
}

int download_host(msg_ctx *c, double *dst, const void *src, size_t cnt) {
  if (c->dtype == MSG_F64) {
    HIPCHK(hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(cnt);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, cnt * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < cnt; ++i) dst[i] = tmp[i];
  }
  return MSG_OK;
}

template <typename S>
int set_image(msg_ctx *c, void **lev, int pair, const double *img) {
  const size_t nn = (size_t)1 << (2 * c->L);
  CHK(upload_host(c, (char *)lev[c->L] + pair * nn * c->es, img, nn));
  return rebuild_pyramid<S>(c, lev, pair, 1);
}

// Relative error bound of one FP32-arithmetic SSD energy evaluation against FP64 arithmetic on the
// same FP32 fields (with a margin over the largest difference measured on STEM-like frames of
// 64^2 to 1024^2 and displacements of up to 6 pixels). Step decisions whose margin lies within
// it are taken on FP64-arithmetic values (needs_exact). FP64 fields and the FP64-arithmetic data
// terms never need it.
constexpr double F32_PASS_TOL = 2e-8;
template <typename S>
double exact_tol_for(const msg_ctx *c) {
  return (std::is_same<S, float>::value && c->data_term == msg::DT_SSD && c->exact_decisions) ? F32_PASS_TOL : 0.0;
}

// ------------------------------------------------------------------------------------------------
// Batched level solve (GradientDescentBase::applySingle for every pair at once)
// ------------------------------------------------------------------------------------------------
template <typename S>
int solve_level_impl(msg_ctx *c, int l, double lambda, int max_iter, double stop_eps, int spec, double *energies,
                     int *iterations, msg_iter_trace *trace, int trace_cap) {
  const int n = 1 << l, B = c->B;
  hipStream_t s = c->stream;
  const bool cons = c->data_term == msg::DT_CONS;  // the consistency energy reads the deformations
  const S *R = cons ? (const S *)c->psi : (const S *)c->R[l], *T = cons ? (const S *)c->psi : (const S *)c->T[l];
  const int npsi = c->npsi;
  S *phi = (S *)c->phi[l];
  S *DE = (S *)c->DE, *z = (S *)c->z, *d = (S *)c->d;
  const MGBufs<S> mg = ctx_mg<S>(c);
  const EnergyGeom eg = energy_geom(n);
  const int nblk_norm = mg_norm_blocks<S>(l, mgc_maxl<S>());
  const int nblk_upd = cdiv(2 * n * n, 256 * msg::UPD_PER_THREAD);
  if (spec <= 0) spec = 3;
  if (spec > msg::MAX_CAND) spec = msg::MAX_CAND;

  msg::IterTrace *tr = nullptr;
  int tcap = 0;
  if (trace && trace_cap > 0) {
    if (c->trace_cap < trace_cap * B) {
      if (c->trace_d) HIPCHK(hipFree(c->trace_d));
      HIPCHK(hipMalloc(&c->trace_d, sizeof(msg::IterTrace) * trace_cap * B));
      c->trace_cap = trace_cap * B;
    }
    tr = c->trace_d;
    tcap = trace_cap;
    HIPCHK(hipMemsetAsync(tr, 0, sizeof(msg::IterTrace) * trace_cap * B, s));
  }

  msg::CtrlArgs a;
  a.nblk_energy = energy_blocks<S>(n, c->data_term);
  a.nblk_norm = nblk_norm;
  a.lambda = lambda;
  a.stop_eps = stop_eps;
  a.max_iter = max_iter;
  a.spec = spec;
  a.trace_cap = tcap;
  a.data_term = c->data_term;
  a.npairs = B;
  a.sigma_fixed = c->sigma_fixed;
  a.sigma0 = c->sigma0;
  a.grad_stop = c->grad_stop;
  a.exact_tol = exact_tol_for<S>(c);

  hipLaunchKernelGGL(k_state_init, dim3(cdiv(B, 64)), dim3(64), 0, s, c->st, B, lambda);
  HIPCHK(hipGetLastError());
  if (c->data_term == msg::DT_NCC)
    CHK(ncc_reference_stats<S>(R, n, c->st, c->part_e, c->part_n, B, a, c->flags_d, tr, s));
  CHK(launch_energy<S>(R, T, phi, z, d, n, c->st, c->part_e, B, s, c->data_term, 0.0, npsi, true, c->yreg));
  a.phase = msg::PH_LEVEL_INIT;
  a.slot = 0;
  hipLaunchKernelGGL(msg::k_control, dim3(B), dim3(msg::CTRL_THREADS), 0, s, c->st, c->part_e, c->part_n, a, c->flags_d, tr);
  HIPCHK(hipGetLastError());

  bool any_active = max_iter > 0;
  int nactive = max_iter > 0 ? B : 0;
  for (int it = 0; any_active; ++it) {
    a.slot = it & 1;
    CHK(launch_grad<S>(R, T, phi, DE, n, c->st, lambda, B, s, c->data_term, npsi, c->yreg));
    CHK(launch_vcycle<S>(DE, 1, z, l, mg, c->st, 1.0, 2 * B, c->part_n, d, s));
    a.phase = msg::PH_ARMIJO_INIT;
    hipLaunchKernelGGL(msg::k_control, dim3(B), dim3(msg::CTRL_THREADS), 0, s, c->st, c->part_e, c->part_n, a, c->flags_d, tr);
    HIPCHK(hipGetLastError());
    bool searching = true;
    int rounds = 0;
    int nsearch = nactive;  // pairs evaluating in this round (upper bound for the first round)
    while (searching) {
      CHK(launch_energy<S>(R, T, phi, z, d, n, c->st, c->part_e, B, s, c->data_term, 0.0, npsi, true, c->yreg));
      a.phase = msg::PH_ARMIJO_STEP;
      hipLaunchKernelGGL(msg::k_control, dim3(B), dim3(msg::CTRL_THREADS), 0, s, c->st, c->part_e, c->part_n, a, c->flags_d,
                         tr);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(c->flags_h, c->flags_d, sizeof(int) * B, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      searching = false;
      any_active = false;
      nsearch = 0;
      nactive = 0;
      for (int p = 0; p < B; ++p) {
        nsearch += (c->flags_h[p] & 1) != 0;
        nactive += (c->flags_h[p] & 2) != 0;
      }
      searching = nsearch > 0;
      any_active = nactive > 0;
      if (++rounds > 200) return fail(MSG_ERR_STATE, "Armijo search did not terminate");
    }
    hipLaunchKernelGGL((msg::k_update<S>), dim3(nblk_upd * B), dim3(256), 0, s, phi, d, z, n, c->st, nblk_upd, B,
                       a.slot);
    HIPCHK(hipGetLastError());
  }

  std::vector<PairState> hs(B);
  HIPCHK(hipMemcpyAsync(hs.data(), c->st, sizeof(PairState) * B, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int p = 0; p < B; ++p) {
    if (energies) energies[p] = hs[p].energy;
    if (iterations) iterations[p] = hs[p].iter;
    c->cnt_iter += hs[p].iter;
    c->cnt_evals += hs[p].total_evals;
    c->cnt_passes += hs[p].total_passes;
    c->cnt_node_iter += (double)hs[p].iter * n * n;
    c->cnt_node_evals += (double)hs[p].total_evals * n * n;
  }
  // algorithmic bytes of the level's energy passes: the first reads R, T and phi, every later one
  // also the direction (the device counts the passes)
  double ebytes = 0;
  for (int p = 0; p < B; ++p) ebytes += (4.0 + 6.0 * hs[p].total_passes) * (double)n * n * sizeof(S);
  prof_add_bytes(KC_ENERGY, ebytes);
  c->level_start.resize(2 * B);
  for (int p = 0; p < B; ++p) {
    c->level_start[2 * p] = hs[p].energy_start;
    c->level_start[2 * p + 1] = hs[p].ereg_start * lambda;
  }
  if (tr) {
    static_assert(sizeof(msg::IterTrace) == sizeof(msg_iter_trace), "trace layout");
    HIPCHK(hipMemcpyAsync(trace, tr, sizeof(msg::IterTrace) * trace_cap * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return MSG_OK;
}

// ------------------------------------------------------------------------------------------------
// Device-driven level solve: every control phase is a one-workgroup-per-pair kernel after the pass
// whose partial sums it reduces, so one iteration is a fixed launch sequence (variation, V-cycle,
// control, CG direction, control, R energy passes each with its control, update) with no host round
// trip. The host replays K iterations
// (as one hipGraph in mode 2) and then reads the per-pair flags once. A pair whose step search is
// still open after the R passes simply sits out the next iteration's gradient work and finishes the
// search in its energy passes, so the result is that of the host-synchronised loop.
// ------------------------------------------------------------------------------------------------
template <typename S>
int solve_level_dd(msg_ctx *c, int l, double lambda, int max_iter, double stop_eps, int spec, double *energies,
                   int *iterations, msg_iter_trace *trace, int trace_cap) {
  const int n = 1 << l, B = c->B;
  hipStream_t s = c->stream;
  const bool cons = c->data_term == msg::DT_CONS;  // the consistency energy reads the deformations
  const S *R = cons ? (const S *)c->psi : (const S *)c->R[l], *T = cons ? (const S *)c->psi : (const S *)c->T[l];
  const int npsi = c->npsi;
  S *phi = (S *)c->phi[l];
  S *DE = (S *)c->DE, *z = (S *)c->z, *d = (S *)c->d;
  const MGBufs<S> mg = ctx_mg<S>(c);
  const EnergyGeom eg = energy_geom(n);
  const int nblk_upd = cdiv(2 * n * n, 256 * msg::UPD_PER_THREAD);
  if (spec <= 0) spec = 3;
  if (spec > msg::MAX_CAND) spec = msg::MAX_CAND;
  const int Rr = 2, K = 8;
  // event timing cannot bracket launches inside a graph: a profiled energy class runs uncaptured
  const bool use_graph = (c->mode == 2) && !((g_prof.mask >> KC_ENERGY) & 1u);

  msg::IterTrace *tr = nullptr;
  int tcap = 0;
  if (trace && trace_cap > 0) {
    if (c->trace_cap < trace_cap * B) {
      if (c->trace_d) HIPCHK(hipFree(c->trace_d));
      HIPCHK(hipMalloc(&c->trace_d, sizeof(msg::IterTrace) * trace_cap * B));
      c->trace_cap = trace_cap * B;
      for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
      c->graphs.clear();
    }
    tr = c->trace_d;
    tcap = trace_cap;
    HIPCHK(hipMemsetAsync(tr, 0, sizeof(msg::IterTrace) * trace_cap * B, s));
  }
  msg::CtrlArgs a;
  a.nblk_energy = energy_blocks<S>(n, c->data_term);
  a.nblk_norm = mg_norm_blocks<S>(l, mgc_maxl<S>());
  a.lambda = lambda;
  a.stop_eps = stop_eps;
  a.max_iter = max_iter;
  a.spec = spec;
  a.trace_cap = tcap;
  a.data_term = c->data_term;
  a.npairs = B;
  a.sigma_fixed = c->sigma_fixed;
  a.sigma0 = c->sigma0;
  a.grad_stop = c->grad_stop;
  a.exact_tol = exact_tol_for<S>(c);
  a.slot = 0;
  // Each control phase is its own one-workgroup-per-pair kernel after the pass whose partial sums it
  // reduces: kernel boundaries make the partials visible across the XCDs (a last-workgroup tail would
  // need an agent-scope release by every workgroup, an L2 write-back on this part, which costs far
  // more than the extra launch).
  auto control = [&](int phase, int slot) -> int {
    msg::CtrlArgs ap = a;
    ap.phase = phase;
    ap.slot = slot;
    ProfScope ps(KC_CONTROL, s, 0);
    hipLaunchKernelGGL(msg::k_control, dim3(B), dim3(msg::CTRL_THREADS), 0, s, c->st, c->part_e, c->part_n, ap,
                       c->flags_d, tr);
    HIPCHK(hipGetLastError());
    return MSG_OK;
  };

  hipLaunchKernelGGL(k_state_init, dim3(cdiv(B, 64)), dim3(64), 0, s, c->st, B, lambda);
  HIPCHK(hipGetLastError());
  if (c->data_term == msg::DT_NCC)
    CHK(ncc_reference_stats<S>(R, n, c->st, c->part_e, c->part_n, B, a, c->flags_d, tr, s));
  CHK(launch_energy<S>(R, T, phi, z, d, n, c->st, c->part_e, B, s, c->data_term, 0.0, npsi, true, c->yreg));
  CHK(control(msg::PH_LEVEL_INIT, 0));

  // iteration k of a replay uses step slot k & 1 (K is even, so the slots keep alternating)
  auto iteration = [&](int k) -> int {
    const int slot = k & 1;
    CHK(launch_grad<S>(R, T, phi, DE, n, c->st, lambda, B, s, c->data_term, npsi, c->yreg));
    CHK(launch_vcycle<S>(DE, 1, z, l, mg, c->st, 1.0, 2 * B, c->part_n, d, s));
    CHK(control(msg::PH_ARMIJO_INIT, slot));
    for (int r = 0; r < Rr; ++r) {
      CHK(launch_energy<S>(R, T, phi, z, d, n, c->st, c->part_e, B, s, c->data_term, 0.0, npsi, true, c->yreg));
      CHK(control(msg::PH_ARMIJO_STEP, slot));
    }
    {
      ProfScope ps(KC_UPDATE, s, 0);
      hipLaunchKernelGGL((msg::k_update<S>), dim3(nblk_upd * B), dim3(256), 0, s, phi, d, z, n, c->st, nblk_upd, B,
                         slot);
      HIPCHK(hipGetLastError());
    }
    return MSG_OK;
  };

  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    for (auto &g : c->graphs)
      if (g.level == l && g.spec == spec && g.K == K && g.R == Rr && g.lambda == lambda && g.stop_eps == stop_eps &&
          g.max_iter == max_iter && g.trace == (void *)tr && g.trace_cap == tcap)
        exec = g.exec;
    if (!exec) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      g_capturing = true;
      int rc = MSG_OK;
      for (int k = 0; k < K && rc == MSG_OK; ++k) rc = iteration(k);
      g_capturing = false;
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc != MSG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIPCHK(ec);
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(ei);
      msg_ctx::GraphEntry ge{l, spec, K, Rr, tcap, lambda, stop_eps, max_iter, (void *)tr, exec};
      c->graphs.push_back(ge);
    }
  }

  // Every pair needs at most max_iter iterations; a search that spills over costs extra launches
  // only. The bound below is far above any legal run and only guards against a logic error.
  // Replays are issued one ahead while a pair can still need the next one (fewer than max_iter
  // iterations replayed), so the stream does not wait for the host to read the flags between
  // replays; each replay's flags are copied to their own snapshot. A replay that turns out not to
  // be needed finds every pair inactive and changes nothing (its kernels return at once).
  const long long max_batches = (long long)(max_iter + 4) * 8 / K + 64;
  auto replay = [&](long long b) -> int {
    if (use_graph) {
      HIPCHK(hipGraphLaunch(exec, s));
    } else {
      for (int k = 0; k < K; ++k) CHK(iteration(k));
    }
    HIPCHK(hipMemcpyAsync(c->flags_h + (b & 1) * B, c->flags_d, sizeof(int) * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->flag_ev[b & 1], s));
    return MSG_OK;
  };
  if (max_iter > 0) {
    CHK(replay(0));
    long long issued = 1;
    for (long long batch = 0;; ++batch) {
      if (batch > max_batches) return fail(MSG_ERR_STATE, "device-driven level solve did not terminate");
      if (issued == batch + 1 && (batch + 1) * K < max_iter) {
        CHK(replay(batch + 1));
        ++issued;
      }
      HIPCHK(hipEventSynchronize(c->flag_ev[batch & 1]));
      bool busy = false;
      for (int p = 0; p < B; ++p) busy |= (c->flags_h[(batch & 1) * B + p] & 3) != 0;
      if (!busy) break;
      if (issued == batch + 1) {
        CHK(replay(batch + 1));
        ++issued;
      }
    }
  }

  std::vector<PairState> hs(B);
  HIPCHK(hipMemcpyAsync(hs.data(), c->st, sizeof(PairState) * B, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int p = 0; p < B; ++p) {
    if (energies) energies[p] = hs[p].energy;
    if (iterations) iterations[p] = hs[p].iter;
    c->cnt_iter += hs[p].iter;
    c->cnt_evals += hs[p].total_evals;
    c->cnt_passes += hs[p].total_passes;
    c->cnt_node_iter += (double)hs[p].iter * n * n;
    c->cnt_node_evals += (double)hs[p].total_evals * n * n;
  }
  // algorithmic bytes of the level's energy passes: the first reads R, T and phi, every later one
  // also the direction (the device counts the passes)
  double ebytes = 0;
  for (int p = 0; p < B; ++p) ebytes += (4.0 + 6.0 * hs[p].total_passes) * (double)n * n * sizeof(S);
  prof_add_bytes(KC_ENERGY, ebytes);
  c->level_start.resize(2 * B);
  for (int p = 0; p < B; ++p) {
    c->level_start[2 * p] = hs[p].energy_start;
    c->level_start[2 * p + 1] = hs[p].ereg_start * lambda;
  }
  if (tr) {
    HIPCHK(hipMemcpyAsync(trace, tr, sizeof(msg::IterTrace) * trace_cap * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return MSG_OK;
}

template <typename S>
int solve_level_any(msg_ctx *c, int l, double lambda, int max_iter, double stop_eps, int spec, double *energies,
                    int *iterations, msg_iter_trace *trace, int trace_cap) {
  if (c->mode == 0)
    return solve_level_impl<S>(c, l, lambda, max_iter, stop_eps, spec, energies, iterations, trace, trace_cap);
  return solve_level_dd<S>(c, l, lambda, max_iter, stop_eps, spec, energies, iterations, trace, trace_cap);
}

template <typename S>
int prolong_deformation(msg_ctx *c, int l, int resample) {
  const int nc = 1 << l, n = 2 * nc;
  if (resample) {
    double *coord = nullptr;
    CHK(resample_coords(nc, n, &coord, c->stream));
    CHK(launch_resample<S>((const S *)c->phi[l], (S *)c->phi[l + 1], nc, n, 2 * c->B, coord, c->stream));
  } else {
    CHK(launch_prolong<S>((const S *)c->phi[l], (S *)c->phi[l + 1], nc, 2 * c->B, c->stream));
  }
  return MSG_OK;
}

// Where the per-level records of a context's solves go (msg_series_levels): the context's pair k
// at level l is record rec0 + k * (stop - start + 1) + (l - start); its field starts at double
// offset field0 + k * per_solve + sum_{start <= m < l} 2 * 4^m.
struct LevelOut {
  const msg_series_levels *lv = nullptr;
  long long rec0 = 0;
  size_t field0 = 0;
};

size_t level_block(int start, int stop) {
  size_t t = 0;
  for (int l = start; l <= stop; ++l) t += (size_t)2 << (2 * l);
  return t;
}

template <typename S>
int record_level(msg_ctx *c, int level, int start, int stop, int max_iter, const LevelOut &o,
                 const std::vector<msg_iter_trace> &tr, const std::vector<int> &its) {
  const msg_series_levels *lv = o.lv;
  const int nl = stop - start + 1;
  const size_t nn2 = (size_t)2 << (2 * level);
  const size_t per_solve = level_block(start, stop);
  const size_t lev_off = level_block(start, level) - nn2;
  std::vector<S> h;
  if (lv->fields || lv->on_level) {
    h.resize(nn2 * c->B);
    HIPCHK(hipMemcpyAsync(h.data(), c->phi[level], sizeof(S) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  std::vector<double> one(lv->on_level ? nn2 : 0);
  for (int k = 0; k < c->B; ++k) {
    const long long r = o.rec0 + (long long)k * nl + (level - start);
    if (lv->fields) {
      double *dst = lv->fields + o.field0 + k * per_solve + lev_off;
      for (size_t i = 0; i < nn2; ++i) dst[i] = (double)h[k * nn2 + i];
    }
    if (lv->on_level) {
      for (size_t i = 0; i < nn2; ++i) one[i] = (double)h[k * nn2 + i];
      const double st2[2] = {c->level_start[2 * k], c->level_start[2 * k + 1]};
      lv->on_level(lv->user, r, level, one.data(), max_iter > 0 ? tr.data() + (size_t)k * max_iter : nullptr,
                   std::min(its[k], std::max(max_iter, 0)), st2);
    }
    if (lv->iters && max_iter > 0)
      std::memcpy(lv->iters + r * max_iter, tr.data() + (size_t)k * max_iter, sizeof(msg_iter_trace) * max_iter);
    if (lv->start) {
      lv->start[2 * r] = c->level_start[2 * k];
      lv->start[2 * r + 1] = c->level_start[2 * k + 1];
    }
    if (lv->iterations) lv->iterations[r] = its[k];
  }
  return MSG_OK;
}

template <typename S>
int solve_impl(msg_ctx *c, int start, int stop, const msg_solve_params *p, double *energies,
               const LevelOut *out = nullptr) {
  std::vector<double> e(c->B, NAN);
  const bool rec = out && out->lv;
  const int cap = rec && (out->lv->iters || out->lv->on_level) ? p->max_iter : 0;
  std::vector<msg_iter_trace> tr((size_t)c->B * std::max(cap, 0));
  std::vector<int> its(c->B, 0);
  for (int level = start; level <= stop; ++level) {
    const double f = p->lambda_factor > 0 ? (double)std::pow(p->lambda_factor, p->lambda_ref_level - level) : 1;
    const double lam = f * p->lambda;
    CHK(solve_level_any<S>(c, level, lam, p->max_iter, p->stop_eps, p->spec, e.data(), its.data(),
                           cap > 0 ? tr.data() : nullptr, cap));
    if (rec) CHK(record_level<S>(c, level, start, stop, cap, *out, tr, its));
    if (level < stop) CHK(prolong_deformation<S>(c, level, p->resample_prolong));
  }
  for (int l = stop; l < c->L; ++l) CHK(prolong_deformation<S>(c, l, p->resample_prolong));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (energies)
    for (int i = 0; i < c->B; ++i) energies[i] = e[i];
  return MSG_OK;
}

}  // namespace

// =================================================================================================
// C ABI
// =================================================================================================
// Every pair evaluates nc step sizes tau0 * 2^-k along its CG direction in the next energy pass.
__global__ void k_state_bench(PairState *st, int npairs, int nc, double tau0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  st[p].dirmode = msg::DIR_CG;
  st[p].ncand = nc;
  st[p].exact = 0;
  for (int k = 0; k < nc; ++k) st[p].cand[k] = tau0 / (double)(1 << k);
}

template <typename S>
int time_energy_pass(msg_ctx *c, int l, int nc, int reps, double *ms, double *bytes) {
  const int n = 1 << l, B = c->B;
  hipStream_t s = c->stream;
  const bool cons = c->data_term == msg::DT_CONS;
  const S *R = cons ? (const S *)c->psi : (const S *)c->R[l], *T = cons ? (const S *)c->psi : (const S *)c->T[l];
  hipLaunchKernelGGL(k_state_bench, dim3(cdiv(B, 64)), dim3(64), 0, s, c->st, B, nc, 1.0);
  HIPCHK(hipGetLastError());
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  int rc = MSG_OK;
  for (int r = 0; r < reps && rc == MSG_OK; ++r)
    rc = launch_energy<S>(R, T, (const S *)c->phi[l], (const S *)c->z, (const S *)c->d, n, c->st, c->part_e, B, s,
                          c->data_term, 0.0, c->npsi, false, c->yreg);
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CHK(rc);
  if (ms) *ms = (double)t / reps;
  if (bytes) *bytes = 6.0 * (double)n * n * sizeof(S) * B;
  return MSG_OK;
}

extern "C" {

const char *msg_last_error(void) { return g_err.c_str(); }
const char *msg_version(void) { return "msgpu 0.1 (gfx950)"; }

int msg_ctx_create(msg_ctx **out, int max_level, int npairs, int dtype) {
  if (!out) return fail(MSG_ERR_ARG, "null context pointer");
  *out = nullptr;
  CHK(check_dtype(dtype));
  if (max_level < 1 || max_level > MAXL) return fail(MSG_ERR_ARG, "max_level out of range");
  if (npairs < 1) return fail(MSG_ERR_ARG, "npairs must be positive");
  msg_ctx *c = new msg_ctx();
  c->mode = g_default_mode;
  c->exact_decisions = g_default_exact;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(MSG_ERR_HIP, "no HIP device");
  }
  c->L = max_level;
  c->B = npairs;
  c->dtype = dtype;
  c->es = esz(dtype);
  const size_t B = npairs;
  int rc = MSG_OK;
  auto A = [&](void **p, size_t bytes) {
    if (rc == MSG_OK) rc = ctx_alloc(c, p, bytes);
  };
  for (int l = 0; l <= max_level; ++l) {
    const size_t nn = (size_t)1 << (2 * l);
    A(&c->R[l], B * nn * c->es);
    A(&c->T[l], B * nn * c->es);
    A(&c->phi[l], B * 2 * nn * c->es);
    if (l < max_level && l <= MAXL) {
      A(&c->mg_rhs[l], B * 2 * nn * c->es);
      A(&c->mg_x2[l], B * 2 * nn * c->es);
      A(&c->mg_x[l], B * 2 * nn * c->es);
    }
  }
  const size_t NN = (size_t)1 << (2 * max_level);
  A(&c->DE, B * 2 * NN * c->es);
  A(&c->z, B * 2 * NN * c->es);
  A(&c->d, B * 2 * NN * c->es);
  A(&c->x2, B * 2 * NN * c->es);
  const int N = 1 << max_level;
  const EnergyGeom eg = energy_geom(N);
  (void)eg;
  A((void **)&c->part_e, sizeof(double) * energy_part_size(N, (int)B));
  A((void **)&c->part_n, sizeof(double) * B * 2 * msg::NORM_TERMS * std::max(1, mg_norm_blocks(max_level, msg::MGC_MAXL)));
  A((void **)&c->st, sizeof(PairState) * B);
  A((void **)&c->flags_d, sizeof(int) * B);
  if (rc != MSG_OK) {
    msg_ctx_destroy(c);
    return rc;
  }
  if (hipHostMalloc((void **)&c->flags_h, sizeof(int) * 2 * B, hipHostMallocDefault) != hipSuccess) {
    msg_ctx_destroy(c);
    return fail(MSG_ERR_ALLOC, "pinned allocation failed");
  }
  for (auto &e : c->flag_ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      msg_ctx_destroy(c);
      return fail(MSG_ERR_HIP, "event creation failed");
    }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    msg_ctx_destroy(c);
    return fail(MSG_ERR_HIP, "stream creation failed");
  }
  // zero images and deformations so that a context is usable before it is filled
  for (int l = 0; l <= max_level; ++l) {
    const size_t nn = (size_t)1 << (2 * l);
    (void)hipMemsetAsync(c->R[l], 0, B * nn * c->es, c->stream);
    (void)hipMemsetAsync(c->T[l], 0, B * nn * c->es, c->stream);
    (void)hipMemsetAsync(c->phi[l], 0, B * 2 * nn * c->es, c->stream);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    msg_ctx_destroy(c);
    return fail(MSG_ERR_HIP, "context initialisation failed");
  }
  *out = c;
  return MSG_OK;
}

int msg_ctx_destroy(msg_ctx *c) {
  if (!c) return MSG_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->trace_d) (void)hipFree(c->trace_d);
  if (c->flags_h) (void)hipHostFree(c->flags_h);
  for (auto &e : c->flag_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MSG_OK;
}

#define DISPATCH(c, fn, ...) ((c)->dtype == MSG_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

static int check_pair(msg_ctx *c, int pair) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (pair < 0 || pair >= c->B) return fail(MSG_ERR_ARG, "pair index out of range");
  return MSG_OK;
}

int msg_set_reference(msg_ctx *c, int pair, const double *img) {
  CHK(check_pair(c, pair));
  if (!img) return fail(MSG_ERR_ARG, "null image");
  return DISPATCH(c, set_image, c, c->R, pair, img);
}

int msg_set_template(msg_ctx *c, int pair, const double *img) {
  CHK(check_pair(c, pair));
  if (!img) return fail(MSG_ERR_ARG, "null image");
  return DISPATCH(c, set_image, c, c->T, pair, img);
}

int msg_set_images_device(msg_ctx *c, const void *d_ref, const void *d_tpl, int dtype) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  CHK(check_dtype(dtype));
  const size_t cnt = ((size_t)1 << (2 * c->L)) * c->B;
  if (d_ref) {
    CHK(convert(d_ref, dtype, c->R[c->L], c->dtype, cnt, c->stream));
    CHK(DISPATCH(c, rebuild_pyramid, c, c->R, -1, 1));
  }
  if (d_tpl) {
    CHK(convert(d_tpl, dtype, c->T[c->L], c->dtype, cnt, c->stream));
    CHK(DISPATCH(c, rebuild_pyramid, c, c->T, -1, 1));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_set_deformation(msg_ctx *c, int pair, const double *phi) {
  CHK(check_pair(c, pair));
  if (!phi) return fail(MSG_ERR_ARG, "null field");
  const size_t nn = (size_t)1 << (2 * c->L);
  CHK(upload_host(c, (char *)c->phi[c->L] + pair * 2 * nn * c->es, phi, 2 * nn));
  CHK(DISPATCH(c, rebuild_pyramid, c, c->phi, pair, 2));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_get_deformation(msg_ctx *c, int pair, double *phi) {
  CHK(check_pair(c, pair));
  if (!phi) return fail(MSG_ERR_ARG, "null field");
  const size_t nn = (size_t)1 << (2 * c->L);
  return download_host(c, phi, (char *)c->phi[c->L] + pair * 2 * nn * c->es, 2 * nn);
}

int msg_set_deformation_level(msg_ctx *c, int pair, int level, const double *phi) {
  CHK(check_pair(c, pair));
  if (!phi) return fail(MSG_ERR_ARG, "null field");
  if (level < 0 || level > c->L) return fail(MSG_ERR_ARG, "level outside 0..max_level");
  const size_t nn = (size_t)1 << (2 * level);
  CHK(upload_host(c, (char *)c->phi[level] + pair * 2 * nn * c->es, phi, 2 * nn));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_get_deformation_level(msg_ctx *c, int pair, int level, double *phi) {
  CHK(check_pair(c, pair));
  if (!phi) return fail(MSG_ERR_ARG, "null field");
  if (level < 0 || level > c->L) return fail(MSG_ERR_ARG, "level outside 0..max_level");
  const size_t nn = (size_t)1 << (2 * level);
  return download_host(c, phi, (char *)c->phi[level] + pair * 2 * nn * c->es, 2 * nn);
}

int msg_set_deformation_zero(msg_ctx *c, int pair) {
  CHK(check_pair(c, pair));
  for (int l = 0; l <= c->L; ++l) {
    const size_t nn = (size_t)1 << (2 * l);
    HIPCHK(hipMemsetAsync((char *)c->phi[l] + pair * 2 * nn * c->es, 0, 2 * nn * c->es, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_set_deformations_device(msg_ctx *c, const void *d_phi) {
  if (!c || !d_phi) return fail(MSG_ERR_ARG, "null argument");
  const size_t cnt = ((size_t)2 << (2 * c->L)) * c->B;
  HIPCHK(hipMemcpyAsync(c->phi[c->L], d_phi, cnt * c->es, hipMemcpyDeviceToDevice, c->stream));
  CHK(DISPATCH(c, rebuild_pyramid, c, c->phi, -1, 2));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_get_deformations_device(msg_ctx *c, void *d_phi) {
  if (!c || !d_phi) return fail(MSG_ERR_ARG, "null argument");
  const size_t cnt = ((size_t)2 << (2 * c->L)) * c->B;
  HIPCHK(hipMemcpyAsync(d_phi, c->phi[c->L], cnt * c->es, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_solve_level(msg_ctx *c, int level, double lambda, int max_iter, double stop_eps, int spec, double *energies,
                    int *iterations, msg_iter_trace *trace, int trace_cap) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (level < 1 || level > c->L) return fail(MSG_ERR_ARG, "level out of range");
  c->cnt_iter = c->cnt_evals = c->cnt_passes = 0;
  return DISPATCH(c, solve_level_any, c, level, lambda, max_iter, stop_eps, spec, energies, iterations, trace,
                  trace_cap);
}

int msg_solve(msg_ctx *c, int start, int stop, const msg_solve_params *p, double *energies) {
  if (!c || !p) return fail(MSG_ERR_ARG, "null argument");
  if (start < 1 || stop > c->L || start > stop) return fail(MSG_ERR_ARG, "invalid level range");
  c->cnt_iter = c->cnt_evals = c->cnt_passes = 0;
  return DISPATCH(c, solve_impl, c, start, stop, p, energies);
}

int msg_last_counts(msg_ctx *c, long long *iterations, long long *evals, long long *passes) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (iterations) *iterations = c->cnt_iter;
  if (evals) *evals = c->cnt_evals;
  if (passes) *passes = c->cnt_passes;
  return MSG_OK;
}

int msg_time_energy_pass(msg_ctx *c, int level, int ncand, int reps, double *ms_per_launch,
                         double *bytes_per_launch) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (level < 1 || level > c->L) return fail(MSG_ERR_ARG, "level out of range");
  if (ncand < 1 || ncand > msg::PLAN_CAND || reps < 1) return fail(MSG_ERR_ARG, "ncand must be 1..3 and reps >= 1");
  return DISPATCH(c, time_energy_pass, c, level, ncand, reps, ms_per_launch, bytes_per_launch);
}

int msg_ctx_buffer(msg_ctx *c, int which, int level, void **ptr) {
  if (!c || !ptr) return fail(MSG_ERR_ARG, "null argument");
  if (level < 0 || level > c->L) return fail(MSG_ERR_ARG, "level out of range");
  if (which == 0) *ptr = c->R[level];
  else if (which == 1) *ptr = c->T[level];
  else if (which == 2) *ptr = c->phi[level];
  else return fail(MSG_ERR_ARG, "unknown buffer");
  return MSG_OK;
}

// ---------------------------------------------------------------------------------------------
// Stateless operator entry points
// ---------------------------------------------------------------------------------------------
}  // extern "C"

namespace {
template <typename S>
// Energy pass at phi for every pair, reduced on the host. SSD: e_data = E_SSD. NCC: e_data = E_NCC and,
// if st_out is given, the pair states the NCC variation reads (reference and template statistics,
// E_NCC at phi).
int energy_impl(const void *R, const void *T, const void *phi, int n, int batch, int data_term, double *e_data,
                       double *e_reg, hipStream_t s, double yreg = 0) {
  const int nb = energy_blocks<S>(n, data_term);
  void *stp = nullptr, *part = nullptr;
  CHK(scratch_get(0, sizeof(PairState) * batch, &stp));
  CHK(scratch_get(1, sizeof(double) * energy_part_size(n, batch), &part));
  hipLaunchKernelGGL(k_state_init, dim3(cdiv(batch, 64)), dim3(64), 0, s, (PairState *)stp, batch, 0.0);
  HIPCHK(hipGetLastError());
  CHK(launch_energy<S>((const S *)R, (const S *)T, (const S *)phi, (const S *)phi, (const S *)phi, n,
                       (const PairState *)stp, (double *)part, batch, s, data_term, 0, 0, true, yreg));
  std::vector<double> hp(energy_part_size(n, batch));
  HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  auto sum = [&](size_t off) {
    double a = 0;
    for (int i = 0; i < nb; ++i) a += hp[off + i];
    return a;
  };
  for (int p = 0; p < batch; ++p) {
    // x-shift: the rows' NCC energies averaged over the rows (FELineShiftRegisEnergy::apply)
    const double dscale = data_term == msg::DT_XSHIFT ? 1.0 / n : 1.0;
    if (e_data) e_data[p] = sum(msg::part_index(msg::DT_SSD, p, 0, 0, nb)) * dscale;
    if (e_reg) e_reg[p] = sum(msg::part_index(msg::DT_SSD, p, 0, 1, nb));
  }
  return MSG_OK;
}

// NCC at phi for every pair through the level solve's own statistics passes and control (the pair
// states are what the NCC variation then reads).
template <typename S>
int ncc_energy_impl(const void *R, const void *T, const void *phi, int n, int batch, double *e_data, double *e_reg,
                    hipStream_t s, std::vector<PairState> *st_out = nullptr) {
  const EnergyGeom g = energy_geom(n);
  void *stp = nullptr, *part = nullptr;
  CHK(scratch_get(0, sizeof(PairState) * batch, &stp));
  CHK(scratch_get(1, sizeof(double) * energy_part_size(n, batch), &part));
  PairState *st = (PairState *)stp;
  hipLaunchKernelGGL(k_state_init, dim3(cdiv(batch, 64)), dim3(64), 0, s, st, batch, 0.0);
  HIPCHK(hipGetLastError());
  msg::CtrlArgs a{};
  a.nblk_energy = g.nblk;
  a.max_iter = 1;  // leaves the pairs active: the variation kernel then takes part
  a.stop_eps = 0;
  a.spec = 1;
  a.data_term = msg::DT_NCC;
  a.npairs = batch;
  CHK(ncc_reference_stats<S>((const S *)R, n, st, (double *)part, nullptr, batch, a, nullptr, nullptr, s));
  CHK(launch_energy<S>((const S *)R, (const S *)T, (const S *)phi, (const S *)phi, (const S *)phi, n, st,
                       (double *)part, batch, s, msg::DT_NCC));
  a.phase = msg::PH_LEVEL_INIT;
  hipLaunchKernelGGL(msg::k_control, dim3(batch), dim3(msg::CTRL_THREADS), 0, s, st, (const double *)part,
                     (const double *)nullptr, a, (int *)nullptr, (msg::IterTrace *)nullptr);
  HIPCHK(hipGetLastError());
  std::vector<PairState> hs(batch);
  HIPCHK(hipMemcpyAsync(hs.data(), st, sizeof(PairState) * batch, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int p = 0; p < batch; ++p) {
    if (e_data) e_data[p] = hs[p].cand_D[0];
    if (e_reg) e_reg[p] = hs[p].cand_R[0];
  }
  if (st_out) *st_out = hs;
  return MSG_OK;
}

// DE = DE_NCC + lambda K phi: an energy pass at phi supplies the statistics the variation needs.
template <typename S>
int ncc_gradient_impl(const void *R, const void *T, const void *phi, void *DE, int n, int batch, double lambda,
                      hipStream_t s) {
  std::vector<PairState> hs;
  CHK(ncc_energy_impl<S>(R, T, phi, n, batch, nullptr, nullptr, s, &hs));
  for (auto &q : hs) q.lambda = lambda;
  void *stp = nullptr;
  CHK(scratch_get(2, sizeof(PairState) * batch, &stp));
  HIPCHK(hipMemcpyAsync(stp, hs.data(), sizeof(PairState) * batch, hipMemcpyHostToDevice, s));
  CHK(launch_grad<S>((const S *)R, (const S *)T, (const S *)phi, (S *)DE, n, (const PairState *)stp, lambda, batch, s,
                     msg::DT_NCC));
  HIPCHK(hipStreamSynchronize(s));
  return MSG_OK;
}
}  // namespace

extern "C" {

int msg_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ssd,
               double *e_reg, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!R || !T || !phi || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64 ? energy_impl<double>(R, T, phi, n, batch, msg::DT_SSD, e_ssd, e_reg, strm(stream))
                          : energy_impl<float>(R, T, phi, n, batch, msg::DT_SSD, e_ssd, e_reg, strm(stream));
}

int msg_ncc_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ncc,
                   double *e_reg, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!R || !T || !phi || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64 ? ncc_energy_impl<double>(R, T, phi, n, batch, e_ncc, e_reg, strm(stream))
                          : ncc_energy_impl<float>(R, T, phi, n, batch, e_ncc, e_reg, strm(stream));
}

int msg_ncc_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch,
                     double lambda, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!R || !T || !phi || !DE || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64 ? ncc_gradient_impl<double>(R, T, phi, DE, n, batch, lambda, strm(stream))
                          : ncc_gradient_impl<float>(R, T, phi, DE, n, batch, lambda, strm(stream));
}

int msg_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch, double lambda,
                 void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!R || !T || !phi || !DE || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64)
    CHK(launch_grad<double>((const double *)R, (const double *)T, (const double *)phi, (double *)DE, n, nullptr,
                            lambda, batch, strm(stream), msg::DT_SSD));
  else
    CHK(launch_grad<float>((const float *)R, (const float *)T, (const float *)phi, (float *)DE, n, nullptr, lambda,
                           batch, strm(stream), msg::DT_SSD));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_xshift_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double y_reg_factor,
                      double *e_ncc, double *e_reg, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (n > (1 << 11)) return fail(MSG_ERR_ARG, "x-shift rows are staged in LDS: n <= 2048");
  if (!R || !T || !phi || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64
             ? energy_impl<double>(R, T, phi, n, batch, msg::DT_XSHIFT, e_ncc, e_reg, strm(stream), y_reg_factor)
             : energy_impl<float>(R, T, phi, n, batch, msg::DT_XSHIFT, e_ncc, e_reg, strm(stream), y_reg_factor);
}

int msg_xshift_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch,
                        double lambda, double y_reg_factor, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (n > (1 << 11)) return fail(MSG_ERR_ARG, "x-shift rows are staged in LDS: n <= 2048");
  if (!R || !T || !phi || !DE || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64)
    CHK(launch_grad<double>((const double *)R, (const double *)T, (const double *)phi, (double *)DE, n, nullptr,
                            lambda, batch, strm(stream), msg::DT_XSHIFT, 0, y_reg_factor));
  else
    CHK(launch_grad<float>((const float *)R, (const float *)T, (const float *)phi, (float *)DE, n, nullptr, lambda,
                           batch, strm(stream), msg::DT_XSHIFT, 0, y_reg_factor));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// msg_energy without allocation: the pair states, the energy pass's partial sums and the reduced
// sums live in one file-scope device array. Each device holds its own copy of the array at its own
// address; host code uses that address (hipGetSymbolAddress, taken once per device), never the
// array's name, which on the host is only a stand-in.
// ------------------------------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEV] = {};

// The current device's scratch address (looked up on the first call on that device).
int noalloc_scratch(void **p) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) return fail(MSG_ERR_ARG, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *q = nullptr;
    HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    if (!q) return fail(MSG_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_noalloc_ptr[dev] = q;
  }
  *p = g_noalloc_ptr[dev];
  return MSG_OK;
}

// Bytes of scratch the no-allocation energy pass needs: pair states, the partials of candidate 0
// (terms E_SSD and E_reg of every pair, laid out as the level solve's partials) and 2 sums per pair.
template <typename S>
size_t noalloc_energy_bytes(int n, int batch, size_t *off_part, size_t *off_sum) {
  const int nblk = energy_geom(n, is_f32_ssd<S>(msg::DT_SSD)).nblk;
  const size_t st = ((sizeof(PairState) * (size_t)batch + 255) / 256) * 256;
  const size_t part_doubles = ((size_t)(batch - 1) * msg::MAX_CAND * 2 + 2) * (size_t)nblk;
  const size_t part = ((sizeof(double) * part_doubles + 255) / 256) * 256;
  *off_part = st;
  *off_sum = st + part;
  return st + part + sizeof(double) * 2 * (size_t)batch;
}

// Sums of the partials in block order (the order msg_energy sums them on the host): out[p] = E_SSD,
// out[batch + p] = E_reg of pair p.
__global__ void k_energy_sums(const double *__restrict__ part, int nblk, int batch, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * batch) return;
  const int p = i % batch, term = i / batch;
  const double *src = part + msg::part_index(msg::DT_SSD, p, 0, term, nblk);
  double a = 0;
  for (int b = 0; b < nblk; ++b) a += src[b];
  out[i] = a;
}

template <typename S>
int energy_noalloc_impl(const void *R, const void *T, const void *phi, int n, int batch, double *e_ssd,
                        double *e_reg, hipStream_t s) {
  size_t off_part = 0, off_sum = 0;
  const size_t need = noalloc_energy_bytes<S>(n, batch, &off_part, &off_sum);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(MSG_ERR_ARG, "msg_energy_noalloc: " + std::to_string(batch) + " pairs of " + std::to_string(n) + "^2 need " +
                                 std::to_string(need) + " bytes of scratch, beyond its " +
                                 std::to_string(NOALLOC_SCRATCH_BYTES) + "; use msg_energy");
  void *base = nullptr;
  CHK(noalloc_scratch(&base));
  unsigned char *b = static_cast<unsigned char *>(base);
  PairState *st = reinterpret_cast<PairState *>(b);
  double *part = reinterpret_cast<double *>(b + off_part);
  double *sums = reinterpret_cast<double *>(b + off_sum);
  const EnergyGeom g = energy_geom(n, is_f32_ssd<S>(msg::DT_SSD));
  const double h = 1.0 / (double)(n - 1);
  // every word the pass reads is written first: states by k_state_init, partials by the pass
  hipLaunchKernelGGL(k_state_init, dim3(cdiv(batch, 64)), dim3(64), 0, s, st, batch, 0.0);
  HIPCHK(hipGetLastError());
  if constexpr (std::is_same<S, float>::value) {
    if (n > 16384) return fail(MSG_ERR_ARG, "FP32 fields: the SSD energy pass addresses grids up to 16384^2");
    const float *Rf = (const float *)R, *Tf = (const float *)T, *pf = (const float *)phi;
    if (n <= 4096 && msg::en32_run(n) == 2)
      hipLaunchKernelGGL((msg::k_energy_f32<true, 2>), dim3(g.nblk * batch), dim3(256), 0, s, Rf, Tf, pf, pf, pf, n,
                         h * h, st, part, g.ntx, g.nty, batch);
    else if (n <= 4096)
      hipLaunchKernelGGL((msg::k_energy_f32<true, 4>), dim3(g.nblk * batch), dim3(256), 0, s, Rf, Tf, pf, pf, pf, n,
                         h * h, st, part, g.ntx, g.nty, batch);
    else
      hipLaunchKernelGGL((msg::k_energy_f32<false, 4>), dim3(g.nblk * batch), dim3(256), 0, s, Rf, Tf, pf, pf, pf, n,
                         h * h, st, part, g.ntx, g.nty, batch);
  }
  else
    hipLaunchKernelGGL((msg::k_energy<S>), dim3(g.nblk * batch), dim3(256), 0, s, (const S *)R, (const S *)T,
                       (const S *)phi, (const S *)phi, (const S *)phi, n, h, (int)((n - 1) * h), st, part, g.ntx,
                       g.nty, batch);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_energy_sums, dim3(cdiv(2 * batch, 64)), dim3(64), 0, s, (const double *)part, g.nblk, batch,
                     sums);
  HIPCHK(hipGetLastError());
  if (e_ssd) HIPCHK(hipMemcpyAsync(e_ssd, sums, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
  if (e_reg) HIPCHK(hipMemcpyAsync(e_reg, sums + batch, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSG_OK;
}
}  // namespace

extern "C" {

int msg_energy_noalloc(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ssd,
                       double *e_reg, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!R || !T || !phi || batch < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64 ? energy_noalloc_impl<double>(R, T, phi, n, batch, e_ssd, e_reg, strm(stream))
                          : energy_noalloc_impl<float>(R, T, phi, n, batch, e_ssd, e_reg, strm(stream));
}

}  // extern "C"

namespace {
template <typename S>
int h1_impl(const void *in, void *out, int n, int nfields, double sigma, hipStream_t s) {
  const int l = log2i(n);
  MGBufs<S> b;
  for (int m = 0; m <= MAXL; ++m) b.rhs[m] = b.x2[m] = b.x[m] = nullptr;
  for (int m = 0; m < l; ++m) {
    const size_t cnt = ((size_t)1 << (2 * m)) * nfields;
    void *p1, *p2, *p3;
    CHK(scratch_get(20 + 3 * m, cnt * sizeof(S), &p1));
    CHK(scratch_get(21 + 3 * m, cnt * sizeof(S), &p2));
    CHK(scratch_get(22 + 3 * m, cnt * sizeof(S), &p3));
    b.rhs[m] = (S *)p1;
    b.x2[m] = (S *)p2;
    b.x[m] = (S *)p3;
  }
  void *top, *tmpout;
  CHK(scratch_get(2, (size_t)n * n * nfields * sizeof(S), &top));
  CHK(scratch_get(3, (size_t)n * n * nfields * sizeof(S), &tmpout));
  b.x2_top = (S *)top;
  // write into a temporary so that in == out is allowed
  CHK(launch_vcycle<S>((const S *)in, 0, (S *)tmpout, l, b, nullptr, sigma, nfields, nullptr, (const S *)nullptr, s));
  HIPCHK(hipMemcpyAsync(out, tmpout, (size_t)n * n * nfields * sizeof(S), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSG_OK;
}
}  // namespace

extern "C" {

int msg_h1_smooth(int dtype, const void *in, void *out, int n, int nfields, double sigma, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!in || !out || nfields < 1) return fail(MSG_ERR_ARG, "invalid argument");
  return dtype == MSG_F64 ? h1_impl<double>(in, out, n, nfields, sigma, strm(stream))
                          : h1_impl<float>(in, out, n, nfields, sigma, strm(stream));
}

int msg_warp(int dtype, const void *img, const void *phi, void *out, int n, int nimg, int phi_shared, int mode,
             double ext, int nearest, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!img || !phi || !out || nimg < 1 || mode < 0 || mode > 2) return fail(MSG_ERR_ARG, "invalid argument");
  const size_t cnt = (size_t)n * n * nimg;
  const dim3 grid((unsigned)((cnt + 255) / 256));
  if (dtype == MSG_F64)
    hipLaunchKernelGGL((msg::k_warp<double>), grid, dim3(256), 0, strm(stream), (const double *)img,
                       (const double *)phi, (double *)out, n, nimg, phi_shared, mode, ext, nearest);
  else
    hipLaunchKernelGGL((msg::k_warp<float>), grid, dim3(256), 0, strm(stream), (const float *)img, (const float *)phi,
                       (float *)out, n, nimg, phi_shared, mode, ext, nearest);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_restrict(int dtype, const void *fine, void *coarse, int n, int nfields, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!fine || !coarse || nfields < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64) CHK(launch_restrict<double>((const double *)fine, (double *)coarse, n, nfields, strm(stream)));
  else CHK(launch_restrict<float>((const float *)fine, (float *)coarse, n, nfields, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_prolong(int dtype, const void *coarse, void *fine, int nc, int nfields, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(2 * nc));
  if (!fine || !coarse || nfields < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64) CHK(launch_prolong<double>((const double *)coarse, (double *)fine, nc, nfields, strm(stream)));
  else CHK(launch_prolong<float>((const float *)coarse, (float *)fine, nc, nfields, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_resample(int dtype, const void *src, void *dst, int nc, int n, int nfields, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  CHK(check_n(nc));
  if (!src || !dst || nfields < 1) return fail(MSG_ERR_ARG, "invalid argument");
  double *coord = nullptr;
  CHK(resample_coords(nc, n, &coord, strm(stream)));
  if (dtype == MSG_F64)
    CHK(launch_resample<double>((const double *)src, (double *)dst, nc, n, nfields, coord, strm(stream)));
  else
    CHK(launch_resample<float>((const float *)src, (float *)dst, nc, n, nfields, coord, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

}  // extern "C"

// =================================================================================================
// Composition, statistics, normalisation, series matching
// =================================================================================================
namespace {

template <typename S>
int compose_impl(const void *phi1, const void *phi2, void *result, int n, hipStream_t s) {
  ProfScope ps(KC_COMPOSE, s, 0);
  const int nn = n * n;
  const int nblk = cdiv(nn, 256);
  void *bc, *rhs, *invd, *x, *g, *hb, *h2, *d0, *d1, *def, *part, *parta, *partb, *U, *nsel, *sel_tmp, *stp, *donep;
  CHK(scratch_get(70, sizeof(double) * 2 * nn, &bc));
  CHK(scratch_get(71, sizeof(double) * 2 * nn, &rhs));
  CHK(scratch_get(72, sizeof(double) * nn, &invd));
  CHK(scratch_get(73, sizeof(double) * 2 * nn, &x));
  CHK(scratch_get(74, sizeof(double) * 2 * nn, &g));
  CHK(scratch_get(75, sizeof(double) * 2 * nn, &hb));
  CHK(scratch_get(76, sizeof(double) * 2 * nn, &h2));
  CHK(scratch_get(77, sizeof(double) * 2 * nn, &d0));
  CHK(scratch_get(78, sizeof(double) * 2 * nn, &d1));
  CHK(scratch_get(79, nn, &def));
  CHK(scratch_get(80, sizeof(double) * 6 * nblk, &part));
  CHK(scratch_get(91, sizeof(double) * 2 * nblk, &parta));
  CHK(scratch_get(92, sizeof(double) * 4 * nblk, &partb));
  CHK(scratch_get(93, sizeof(int) * nn, &U));
  CHK(scratch_get(94, sizeof(int) * 4, &nsel));
  CHK(scratch_get(95, sizeof(msg::CgState) * 4, &stp));
  CHK(scratch_get(96, sizeof(int) * 2, &donep));
  msg::CgState *sa = (msg::CgState *)stp, *sb = sa + 2;
  int *done = (int *)donep;
  const dim3 g1(nblk), b(256);
  hipLaunchKernelGGL((msg::k_compose_init<S>), g1, b, 0, s, (const S *)phi1, (const S *)phi2, (double *)bc,
                     (unsigned char *)def, n);
  hipLaunchKernelGGL(msg::k_compose_rhs, dim3(cdiv(2 * nn, 256)), b, 0, s, (const double *)bc,
                     (const unsigned char *)def, (double *)rhs, (double *)invd, n);
  HIPCHK(hipGetLastError());
  // the undefined nodes, in index order
  size_t tbytes = 0;
  const msg::UndefinedNode pred{(const unsigned char *)def};
  HIPCHK(hipcub::DeviceSelect::If(nullptr, tbytes, hipcub::CountingInputIterator<int>(0), (int *)U, (int *)nsel, nn,
                                  pred, s));
  CHK(scratch_get(97, tbytes, &sel_tmp));
  HIPCHK(hipcub::DeviceSelect::If(sel_tmp, tbytes, hipcub::CountingInputIterator<int>(0), (int *)U, (int *)nsel, nn,
                                  pred, s));
  // the start value is the result's previous content (PCGInverse::apply starts from its Dest)
  CHK(convert(result, sizeof(S) == 8 ? MSG_F64 : MSG_F32, x, MSG_F64, 2 * (size_t)nn, s));
  hipLaunchKernelGGL(msg::k_cg_start, dim3(nblk, 2), b, 0, s, (const double *)x, (const double *)rhs,
                     (const unsigned char *)def, (const double *)invd, (double *)g, (double *)hb, (double *)d0,
                     (double *)part, n, nblk);
  hipLaunchKernelGGL(msg::k_cg_init, dim3(2), b, 0, s, (const double *)part, nblk, sb, done);
  HIPCHK(hipGetLastError());
  int nu = 0;
  HIPCHK(hipMemcpyAsync(&nu, nsel, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int nbu = std::max(1, cdiv(nu, 256));
  double *dbuf[2] = {(double *)d0, (double *)d1};
  // Blocks of 8 iterations, each followed by a snapshot of the done flags; the host checks a block's
  // snapshot while the next block runs (iterations past convergence change nothing).
  FlagSnap *fs = nullptr;
  CHK(flag_snap(&fs));
  int k = 0;
  auto block = [&](int blk) -> int {
    for (int r = 0; r < 8 && k <= 1000; ++r, ++k) {
      const double *dp = dbuf[k > 0 ? (k - 1) & 1 : 0];
      double *dc = dbuf[k & 1];
      hipLaunchKernelGGL(msg::k_cg_a, dim3(nbu, 2), b, 0, s, k, (const int *)U, nu, (const unsigned char *)def, dp, dc,
                         (const double *)hb, (double *)h2, (const double *)partb, (double *)parta, nbu,
                         (const msg::CgState *)sb, sa, done, n);
      hipLaunchKernelGGL(msg::k_cg_b, dim3(nbu, 2), b, 0, s, (const int *)U, nu, (double *)x, (const double *)dc,
                         (double *)g, (double *)hb, (const double *)h2, (const double *)invd, (const double *)parta,
                         (double *)partb, nbu, (const msg::CgState *)sa, sb, (const int *)done, n);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(fs->h + 2 * (blk & 1), done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(fs->ev[blk & 1], s));
    return MSG_OK;
  };
  // the state after the start (k_cg_init may already have met the stopping rule)
  HIPCHK(hipMemcpyAsync(fs->h + 2, done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(fs->ev[1], s));
  HIPCHK(hipEventSynchronize(fs->ev[1]));
  if (!(fs->h[2] && fs->h[3])) {
    CHK(block(0));
    for (int blk = 0;; ++blk) {
      const bool more = k <= 1000;
      if (more) CHK(block(blk + 1));
      HIPCHK(hipEventSynchronize(fs->ev[blk & 1]));
      if ((fs->h[2 * (blk & 1)] && fs->h[2 * (blk & 1) + 1]) || !more) break;
    }
  }
  hipLaunchKernelGGL((msg::k_cg_finish<S>), dim3(cdiv(2 * nn, 256)), b, 0, s, (const double *)x, (const double *)rhs,
                     (const unsigned char *)def, (const msg::CgState *)sb, (S *)result, n);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

template <typename S>
int stats_impl(const void *stack, int nimg, int n, double fill, void *mean, void *count, void *median, hipStream_t s) {
  const int nn = n * n;
  const int P = msg::stats_pixels_per_wave(nimg);
  if (P == 0) return fail(MSG_ERR_ARG, "series longer than 8192 frames are not supported by the median kernel");
  const size_t lds = sizeof(double) * P * std::max(nimg, 1);
  const dim3 g(cdiv(nn, P)), b(64);
  const S *st = (const S *)stack;
  if (P == 64) hipLaunchKernelGGL((msg::k_series_stats<S, 64>), g, b, lds, s, st, nimg, nn, fill, (S *)mean, (S *)count, (S *)median);
  else if (P == 16) hipLaunchKernelGGL((msg::k_series_stats<S, 16>), g, b, lds, s, st, nimg, nn, fill, (S *)mean, (S *)count, (S *)median);
  else if (P == 4) hipLaunchKernelGGL((msg::k_series_stats<S, 4>), g, b, lds, s, st, nimg, nn, fill, (S *)mean, (S *)count, (S *)median);
  else hipLaunchKernelGGL((msg::k_series_stats<S, 1>), g, b, lds, s, st, nimg, nn, fill, (S *)mean, (S *)count, (S *)median);
  HIPCHK(hipGetLastError());
  return MSG_OK;
}

template <typename S>
int normalize_impl(void *img, int n, int nimg, double pct, int min_to_zero, hipStream_t s) {
  const int nn = n * n;
  void *sorted, *tmp;
  CHK(scratch_get(81, sizeof(S) * nn, &sorted));
  size_t tbytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tbytes, (const S *)img, (S *)sorted, nn, 0, sizeof(S) * 8, s));
  CHK(scratch_get(82, tbytes, &tmp));
  const int m1 = nn - 1;
  int i0 = 0, i1 = m1;
  if (pct >= 0) {
    i0 = (int)(pct / 200 * m1);
    i1 = (int)((1 - pct / 200) * m1);
    i0 = std::min(std::max(i0, 0), m1);
    i1 = std::min(std::max(i1, 0), m1);
  }
  for (int k = 0; k < nimg; ++k) {
    S *im = (S *)img + (size_t)k * nn;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp, tbytes, (const S *)im, (S *)sorted, nn, 0, sizeof(S) * 8, s));
    hipLaunchKernelGGL((msg::k_normalize<S>), dim3(cdiv(nn, 256)), dim3(256), 0, s, im, nn, (const S *)sorted, i0, i1,
                       pct >= 0 ? 1 : 0, min_to_zero);
    HIPCHK(hipGetLastError());
  }
  return MSG_OK;
}

template <typename S>
__global__ void k_sum_blocks(const S *__restrict__ a, int nn, double *__restrict__ part) {
  __shared__ double red[4];
  double v = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256) v += msg::ld(a, i);
  const double s = msg::block_sum<256>(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename S>
int mean_value(const void *img, int nn, double *out, hipStream_t s) {
  void *part;
  const int nb = 256;
  CHK(scratch_get(83, sizeof(double) * nb, &part));
  hipLaunchKernelGGL((k_sum_blocks<S>), dim3(nb), dim3(256), 0, s, (const S *)img, nn, (double *)part);
  HIPCHK(hipGetLastError());
  double hp[256];
  HIPCHK(hipMemcpyAsync(hp, part, sizeof(hp), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double t = 0;
  for (int i = 0; i < nb; ++i) t += hp[i];
  *out = t / nn;
  return MSG_OK;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Contexts of msg_match_series are kept for reuse by later calls with the same shape (a stage is
// typically run many times on series of one size); each is handed to one call at a time.
struct CtxCache {
  std::mutex mu;
  std::vector<msg_ctx *> free;
};
CtxCache g_ctx_cache;

// Iterations, reference-equivalent energy evaluations and energy passes of every solve run on a
// cached context since the last reset (the work of the last msg_match_series call).
struct SeriesCounts {
  std::mutex mu;
  long long v[3] = {0, 0, 0};
  double w[2] = {0, 0};
};
SeriesCounts g_series_counts;

int ctx_acquire(int L, int B, int dtype, msg_ctx **out) {
  {
    std::lock_guard<std::mutex> g(g_ctx_cache.mu);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    for (size_t i = 0; i < g_ctx_cache.free.size(); ++i) {
      msg_ctx *c = g_ctx_cache.free[i];
      if (c->L == L && c->B == B && c->dtype == dtype && c->device == dev) {
        g_ctx_cache.free.erase(g_ctx_cache.free.begin() + i);
        c->cnt_iter = c->cnt_evals = c->cnt_passes = 0;
        c->cnt_node_iter = c->cnt_node_evals = 0;
        *out = c;
        return MSG_OK;
      }
    }
  }
  return msg_ctx_create(out, L, B, dtype);
}

void ctx_release(msg_ctx *c) {
  if (!c) return;
  {
    std::lock_guard<std::mutex> g(g_series_counts.mu);
    g_series_counts.v[0] += c->cnt_iter;
    g_series_counts.v[1] += c->cnt_evals;
    g_series_counts.v[2] += c->cnt_passes;
    g_series_counts.w[0] += c->cnt_node_iter;
    g_series_counts.w[1] += c->cnt_node_evals;
  }
  std::lock_guard<std::mutex> g(g_ctx_cache.mu);
  g_ctx_cache.free.push_back(c);
}

struct CtxGuard {
  msg_ctx *c;
  ~CtxGuard() { ctx_release(c); }
};

// SeriesMatching::reduceDeformations (matchSeries.h:678-704) on n device fields at level L: a one-pair
// context whose data term is the consistency energy of the n deformations, plain H1GradientDescent
// (filter width sqrt(10 h) fixed, gradientDescent.h:639-650) and gradient-norm stopping.
template <typename S>
int reduce_impl(const void *defs, int n, int L, int max_iter, double stop_eps, void *phi_out, double *energy,
                int *iterations) {
  const int N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = sizeof(S) == 8 ? MSG_F64 : MSG_F32;
  msg_ctx *c = nullptr;
  CHK(ctx_acquire(L, 1, dt, &c));
  CtxGuard g{c};
  auto reset = [c]() {  // captured graphs hold the deformations' address: never keep them
    for (auto &ge : c->graphs) (void)hipGraphExecDestroy(ge.exec);
    c->graphs.clear();
  };
  reset();
  c->data_term = msg::DT_CONS;
  c->psi = defs;
  c->npsi = n;
  c->sigma_fixed = 1;
  c->sigma0 = std::sqrt(10.0 / (double)(N - 1));
  c->grad_stop = 1;
  HIPCHK(hipMemsetAsync(c->phi[L], 0, sizeof(S) * 2 * nn, c->stream));
  double e = 0;
  int it = 0;
  const int rc = solve_level_any<S>(c, L, 0.0, max_iter, stop_eps, 0, &e, &it, nullptr, 0);
  if (rc == MSG_OK && phi_out)
    (void)hipMemcpyAsync(phi_out, c->phi[L], sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, c->stream);
  const hipError_t es = hipStreamSynchronize(c->stream);
  reset();
  c->data_term = msg::DT_SSD;
  c->psi = nullptr;
  c->npsi = 0;
  c->sigma_fixed = c->grad_stop = 0;
  c->sigma0 = 1.0;
  CHK(rc);
  HIPCHK(es);
  if (energy) *energy = e;
  if (iterations) *iterations = it;
  return MSG_OK;
}

// One background host thread for work that overlaps the caller's: msg_match_series runs its
// pairwise batches there while the calling thread runs the chain. Never destroyed (a static
// condition variable with a blocked waiter must not be destroyed at exit).
struct Worker {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  bool started = false;
  void post(std::function<void()> f) {
    std::lock_guard<std::mutex> g(mu);
    if (!started) {
      std::thread([this] { loop(); }).detach();
      started = true;
    }
    q.push_back(std::move(f));
    cv.notify_one();
  }
  void loop() {
    for (;;) {
      std::function<void()> f;
      {
        std::unique_lock<std::mutex> l(mu);
        cv.wait(l, [&] { return !q.empty(); });
        f = std::move(q.front());
        q.pop_front();
      }
      f();
    }
  }
};
Worker *worker() {
  static Worker *w = new Worker;
  return w;
}

// Progress of the pairwise batches as seen by the chain.
struct Progress {
  std::mutex mu;
  std::condition_variable cv;
  int done = 0;           // pairwise solves finished (a prefix of the pairs)
  bool finished = false;  // the batch task has returned
  bool abort = false;     // set by the chain after an error: no further batches
  int rc = MSG_OK;
  std::string err;
};

// Device buffers of the phase correlation of n x n frames, allocated on first use and kept for the
// following pairs of the same size.
struct TranslationWork {
  int n = 0;
  DevBuf w, x, y, z, corr, part, win;
  size_t win_cap = 0;  // doubles held by `win`
  // The (2m+1)^2 correlation values of a maxCorrShift window (m < n, so up to (2n-1)^2 > 2 n^2:
  // a buffer of its own, grown on demand).
  int ensure_window(size_t cnt) {
    if (cnt <= win_cap) return MSG_OK;
    if (win.p) {
      HIPCHK(hipFree(win.p));
      win.p = nullptr;
    }
    win_cap = 0;
    HIPCHK(hipMalloc(&win.p, sizeof(double) * cnt));
    win_cap = cnt;
    return MSG_OK;
  }
  int ensure(int nn_side) {
    if (n == nn_side) return MSG_OK;
    const size_t nn = (size_t)nn_side * nn_side;
    for (DevBuf *b : {&w, &x, &y, &z, &corr, &part})
      if (b->p) {
        HIPCHK(hipFree(b->p));
        b->p = nullptr;
      }
    n = 0;
    HIPCHK(hipMalloc(&w.p, sizeof(double2) * (nn_side / 2)));
    HIPCHK(hipMalloc(&x.p, sizeof(double2) * nn));
    HIPCHK(hipMalloc(&y.p, sizeof(double2) * nn));
    HIPCHK(hipMalloc(&z.p, sizeof(double2) * nn));
    HIPCHK(hipMalloc(&corr.p, sizeof(double) * nn));
    HIPCHK(hipMalloc(&part.p, (sizeof(double) + sizeof(long long)) * 1024));
    n = nn_side;
    return MSG_OK;
  }
};

// PhaseCorrelationRegistration::registerImages(Reference, Template, MaxInfAbsShift)
// (quoc/paramReg.h:2057-2096): FFT2 of both frames, the (normalised) cross power spectrum
// FFT(reference) * conj(FFT(template)), inverse FFT scaled by 1/n^2 (convolution.h:1049-1114), then
// the first maximum, or with max_shift > 0 the strict maximum over the window in x-fastest order
// from (0, 0). The shift is returned signed, (-n/2, n/2].
template <typename S>
int translation_impl(const S *refimg, const S *tplimg, int n, int max_shift, int *shift, double *corr_out,
                     TranslationWork &wk, hipStream_t s) {
  const int logn = log2i(n);
  if (logn > msg::FFT_MAXLOG) return fail(MSG_ERR_ARG, "phase correlation supports frames of at most 4096 x 4096");
  if (max_shift >= n) return fail(MSG_ERR_ARG, "maxCorrShift must be smaller than the frame size");
  CHK(wk.ensure(n));
  const size_t nn = (size_t)n * n;
  double2 *W = (double2 *)wk.w.p, *X = (double2 *)wk.x.p, *Y = (double2 *)wk.y.p, *Z = (double2 *)wk.z.p;
  double *corr = corr_out ? corr_out : (double *)wk.corr.p;
  hipLaunchKernelGGL(msg::k_fft_twiddles, dim3(cdiv(n / 2, 256)), dim3(256), 0, s, W, n);
  const int thr = std::min(1024, std::max(64, n / 2));
  const size_t lds = sizeof(double2) * n;
  const dim3 tg(cdiv(n, 32), cdiv(n, 32)), tb(32, 8);
  const unsigned eg = (unsigned)std::min<size_t>((nn + 255) / 256, 8192);
  // rows, transpose, rows: Y = FFT2(reference)^T, Z = FFT2(template)^T
  hipLaunchKernelGGL((msg::k_fft_rows<S>), dim3(n), dim3(thr), lds, s, refimg, X, W, n, logn, 0);
  hipLaunchKernelGGL((msg::k_fft_rows<S>), dim3(n), dim3(thr), lds, s, tplimg, Y, W, n, logn, 0);
  hipLaunchKernelGGL(msg::k_transpose_c, tg, tb, 0, s, X, Z, n);
  hipLaunchKernelGGL(msg::k_transpose_c, tg, tb, 0, s, Y, X, n);
  hipLaunchKernelGGL((msg::k_fft_rows<double2>), dim3(n), dim3(thr), lds, s, Z, Y, W, n, logn, 0);
  hipLaunchKernelGGL((msg::k_fft_rows<double2>), dim3(n), dim3(thr), lds, s, X, Z, W, n, logn, 0);
  hipLaunchKernelGGL(msg::k_cross_power, dim3(eg), dim3(256), 0, s, Y, Z, nn, max_shift > 0 ? 0 : 1);
  // inverse: rows, transpose, rows, real part / n^2
  hipLaunchKernelGGL((msg::k_fft_rows<double2>), dim3(n), dim3(thr), lds, s, Y, X, W, n, logn, 1);
  hipLaunchKernelGGL(msg::k_transpose_c, tg, tb, 0, s, X, Z, n);
  hipLaunchKernelGGL((msg::k_fft_rows<double2>), dim3(n), dim3(thr), lds, s, Z, Y, W, n, logn, 1);
  hipLaunchKernelGGL(msg::k_real_part, dim3(eg), dim3(256), 0, s, Y, corr, nn, 1.0 / (double)nn);
  HIPCHK(hipGetLastError());
  int sx = 0, sy = 0;
  if (max_shift <= 0) {
    const int nb = (int)std::min<size_t>((nn + 255) / 256, 1024);
    double *bv = (double *)wk.part.p;
    long long *bi = (long long *)(bv + 1024);
    hipLaunchKernelGGL(msg::k_argmax_part, dim3(nb), dim3(256), 0, s, corr, nn, bv, bi);
    HIPCHK(hipGetLastError());
    std::vector<double> hv(nb);
    std::vector<long long> hi(nb);
    HIPCHK(hipMemcpyAsync(hv.data(), bv, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hi.data(), bi, sizeof(long long) * nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double best = 0;
    long long at = -1;
    for (int b = 0; b < nb; ++b)
      if (hi[b] >= 0 && (at < 0 || hv[b] > best || (hv[b] == best && hi[b] < at))) {
        best = hv[b];
        at = hi[b];
      }
    if (at < 0) at = 0;
    sx = (int)(at % n);
    sy = (int)(at / n);
  } else {
    const int side = 2 * max_shift + 1;
    const size_t cnt = (size_t)side * side;
    CHK(wk.ensure_window(cnt));
    double *win = (double *)wk.win.p;
    hipLaunchKernelGGL(msg::k_gather_window, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, corr, n, max_shift,
                       win);
    HIPCHK(hipGetLastError());
    std::vector<double> hw(cnt);
    HIPCHK(hipMemcpyAsync(hw.data(), win, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double best = hw[(size_t)max_shift * side + max_shift];  // the value at shift (0, 0)
    for (int dy = -max_shift; dy <= max_shift; ++dy)
      for (int dx = -max_shift; dx <= max_shift; ++dx) {
        const double v = hw[(size_t)(dy + max_shift) * side + (dx + max_shift)];
        if (v > best) {
          best = v;
          sx = (dx + n) % n;
          sy = (dy + n) % n;
        }
      }
  }
  shift[0] = sx <= n / 2 ? sx : sx - n;
  shift[1] = sy <= n / 2 ? sy : sy - n;
  return MSG_OK;
}

// qc::approxInverseDeformation (deformations.h:341-355): keys of the first covering triangle, the
// interpolated identity minus the identity, the covered range's midpoint where nothing covers.
template <typename S>
int inverse_impl(const S *phi, S *out, int n, hipStream_t s) {
  const size_t nn = (size_t)n * n;
  DevBuf key, part;
  const int nb = (int)std::min<size_t>((nn + 255) / 256, 1024);
  HIPCHK(hipMalloc(&key.p, sizeof(unsigned) * nn));
  HIPCHK(hipMalloc(&part.p, sizeof(double) * 4 * nb));
  unsigned *kp = (unsigned *)key.p;
  const double H = 1.0 / (double)(n - 1);
  const unsigned eg = (unsigned)std::min<size_t>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(msg::k_fill_u32, dim3(eg), dim3(256), 0, s, kp, nn, msg::INV_NONE);
  const long long ne = (long long)(n - 1) * (n - 1);
  hipLaunchKernelGGL((msg::k_inv_keys<S>), dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, phi, n, H, kp);
  hipLaunchKernelGGL((msg::k_inv_values<S>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, phi, n, H,
                     (const unsigned *)kp, out);
  hipLaunchKernelGGL((msg::k_inv_minmax<S>), dim3(nb), dim3(256), 0, s, (const S *)out, (const unsigned *)kp, nn,
                     (double *)part.p);
  HIPCHK(hipGetLastError());
  std::vector<double> hp(4 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double mn[2] = {DBL_MAX, DBL_MAX}, mx[2] = {-DBL_MAX, -DBL_MAX};
  for (int b = 0; b < nb; ++b)
    for (int c = 0; c < 2; ++c) {
      mn[c] = std::min(mn[c], hp[4 * b + 2 * c]);
      mx[c] = std::max(mx[c], hp[4 * b + 2 * c + 1]);
    }
  hipLaunchKernelGGL((msg::k_inv_fill<S>), dim3(eg), dim3(256), 0, s, out, (const unsigned *)kp, nn,
                     0.5 * (mn[0] + mx[0]), 0.5 * (mn[1] + mx[1]));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));  // the key buffer is freed on return
  return MSG_OK;
}

// Pairwise solves i0..i1-1 in one batched context: template i against its predecessor (the
// reference for i = 0 or with dont_accumulate), SeriesMatching::matchAndAverageSeries
// (matchSeries.h:136-176). Writes the fields to pw and the energies.
template <typename S>
int pairwise_batch(const msg_series_params *p, const msg_solve_params &sp, const S *tpl, const S *ref, int i0,
                   int i1, S *pw, double *energies, const msg_series_levels *lv, const S *init0, S *first,
                   long long alt_rec0, size_t alt_field0) {
  const int L = p->precision_level;
  const int N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = sizeof(S) == 8 ? MSG_F64 : MSG_F32;
  const int nb = i1 - i0;
  msg_ctx *pc = nullptr;
  CHK(ctx_acquire(L, nb, dt, &pc));
  CtxGuard gpc{pc};
  CHK(msg_ctx_set_data_term(pc, p->data_term));
  if (p->data_term == MSG_XSHIFT) CHK(msg_ctx_set_y_reg_factor(pc, p->y_reg_factor));
  hipStream_t s = pc->stream;
  for (int i = i0; i < i1; ++i) {
    const S *r = (i == 0 || p->dont_accumulate) ? ref : tpl + (size_t)(i - 1) * nn;
    HIPCHK(hipMemcpyAsync((S *)pc->R[L] + (i - i0) * nn, r, nn * sizeof(S), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync((S *)pc->T[L] + (i - i0) * nn, tpl + (size_t)i * nn, nn * sizeof(S),
                          hipMemcpyDeviceToDevice, s));
  }
  if (!p->dont_normalize) {
    CHK(normalize_impl<S>(pc->R[L], N, nb, p->saturation_pct, p->normalize_min_to_zero, s));
    CHK(normalize_impl<S>(pc->T[L], N, nb, p->saturation_pct, p->normalize_min_to_zero, s));
  }
  if (p->pre_smooth_sigma > 0) {  // prepareMultilevelArray's optional pre-smoothing (registration.h:356-362)
    const double sig = p->pre_smooth_sigma / (double)(N - 1);
    CHK(h1_impl<S>(pc->R[L], pc->R[L], N, nb, sig, s));
    CHK(h1_impl<S>(pc->T[L], pc->T[L], N, nb, sig, s));
  }
  CHK(rebuild_pyramid<S>(pc, pc->R, -1, 1));
  CHK(rebuild_pyramid<S>(pc, pc->T, -1, 1));
  // every pairwise solve starts from the zero deformation (a context taken from the cache still
  // holds its previous solution)
  for (int l = 0; l <= L; ++l)
    HIPCHK(hipMemsetAsync(pc->phi[l], 0, sizeof(S) * ((size_t)2 << (2 * l)) * nb, s));
  if (init0 && i0 == 0) {  // template 0 starts from the given deformation (matchSeries.h:123-135)
    HIPCHK(hipMemcpyAsync(pc->phi[L], init0, sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, s));
    CHK(rebuild_pyramid<S>(pc, pc->phi, 0, 2));
  }
  if (p->corr_init) {
    // "useCorrelationToInitTranslation": the translation found by phase correlation of the
    // template (as registerImages' reference) against the reference, in grid units, on every
    // level (setTransformationToTranslation, registration.h:578-582; matchSeries.h:145-153)
    TranslationWork wk;
    const double h = 1.0 / (double)(N - 1);
    for (int i = i0; i < i1; ++i) {
      const int k = i - i0;
      int sh[2];
      CHK(translation_impl<S>((const S *)pc->T[L] + k * nn, (const S *)pc->R[L] + k * nn, N, p->max_corr_shift, sh,
                              nullptr, wk, s));
      hipLaunchKernelGGL((msg::k_fill_translation<S>), dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)),
                         dim3(256), 0, s, (S *)pc->phi[L] + k * 2 * nn, nn, (double)sh[0] * h, (double)sh[1] * h);
      HIPCHK(hipGetLastError());
      CHK(rebuild_pyramid<S>(pc, pc->phi, k, 2));
    }
  }
  LevelOut lo;
  lo.lv = lv;
  lo.rec0 = (long long)i0 * (p->stop_level - p->start_level + 1);
  lo.field0 = (size_t)i0 * level_block(p->start_level, p->stop_level);
  CHK(solve_impl<S>(pc, p->start_level, p->stop_level, &sp, energies + i0, &lo));
  if (first) HIPCHK(hipMemcpyAsync(first + (size_t)i0 * 2 * nn, pc->phi[L], sizeof(S) * 2 * nn * nb,
                                   hipMemcpyDeviceToDevice, s));
  if (p->alt_start_level > 0 && p->alt_start_level != p->start_level) {
    // altStartLevel (matchSeries.h:158-203): the solve repeated from zero at alt_start_level; per pair
    // the first result is kept if it leaves fewer template pixels outside the domain (DeformImage with
    // +Inf outside), or with energy_as_alt_criterion if its energy is lower
    auto outside = [&](std::vector<long long> &cnt) -> int {
      DevBuf w;
      HIPCHK(hipMalloc(&w.p, sizeof(S) * nn * nb));
      const size_t c = nn * nb;
      hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((c + 255) / 256)), dim3(256), 0, s, (const S *)pc->T[L],
                         (const S *)pc->phi[L], (S *)w.p, N, nb, 0, 0, (double)INFINITY, 0);
      HIPCHK(hipGetLastError());
      std::vector<S> h(c);
      HIPCHK(hipMemcpyAsync(h.data(), w.p, sizeof(S) * c, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      cnt.assign(nb, 0);
      for (size_t k = 0; k < c; ++k) cnt[k / nn] += std::isinf(h[k]) && h[k] > 0;
      return MSG_OK;
    };
    std::vector<long long> out1, out2;
    CHK(outside(out1));
    const std::vector<double> e1(energies + i0, energies + i1);
    DevBuf keep;
    HIPCHK(hipMalloc(&keep.p, sizeof(S) * 2 * nn * nb));
    HIPCHK(hipMemcpyAsync(keep.p, pc->phi[L], sizeof(S) * 2 * nn * nb, hipMemcpyDeviceToDevice, s));
    for (int l = 0; l <= L; ++l)
      HIPCHK(hipMemsetAsync(pc->phi[l], 0, sizeof(S) * ((size_t)2 << (2 * l)) * nb, s));
    const int na = p->stop_level - p->alt_start_level + 1;
    LevelOut la;
    la.lv = lv;
    la.rec0 = alt_rec0 + (long long)i0 * na;
    la.field0 = alt_field0 + (size_t)i0 * level_block(p->alt_start_level, p->stop_level);
    std::vector<double> e2(nb);
    CHK(solve_impl<S>(pc, p->alt_start_level, p->stop_level, &sp, e2.data(), &la));
    CHK(outside(out2));
    for (int k = 0; k < nb; ++k) {
      const bool keepFirst = p->energy_as_alt_criterion ? e1[k] < e2[k] : out1[k] < out2[k];
      if (keepFirst)
        HIPCHK(hipMemcpyAsync((S *)pc->phi[L] + (size_t)k * 2 * nn, (const S *)keep.p + (size_t)k * 2 * nn,
                              sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, s));
      energies[i0 + k] = keepFirst ? e1[k] : e2[k];
    }
  }
  HIPCHK(hipMemcpyAsync(pw + (size_t)i0 * 2 * nn, pc->phi[L], sizeof(S) * 2 * nn * nb, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSG_OK;
}

// Runs the batches [k*b, min((k+1)*b, npw)) in order, publishing progress after each.
template <typename S>
void pairwise_all(const msg_series_params *p, const msg_solve_params &sp, const S *tpl, const S *ref, int npw,
                  int b, S *pw, double *energies, const msg_series_levels *lv, const S *init0, Progress *pr, S *first,
                  long long alt_rec0, size_t alt_field0) {
  int rc = MSG_OK;
  for (int i0 = 0; i0 < npw && rc == MSG_OK; i0 += b) {
    {
      std::lock_guard<std::mutex> g(pr->mu);
      if (pr->abort) break;
    }
    rc = pairwise_batch<S>(p, sp, tpl, ref, i0, std::min(i0 + b, npw), pw, energies, lv, init0, first, alt_rec0,
                           alt_field0);
    std::lock_guard<std::mutex> g(pr->mu);
    if (rc == MSG_OK) pr->done = std::min(i0 + b, npw);
    else { pr->rc = rc; pr->err = g_err; }
    pr->cv.notify_all();
  }
  std::lock_guard<std::mutex> g(pr->mu);
  pr->finished = true;
  pr->cv.notify_all();
}

template <typename S>
int match_series_impl(const msg_series_params *p, const void *templates, int n, const void *reference,
                      const void *pairwise_in, void *deformations, void *pairwise_out, double *energies, void *mean,
                      void *count, void *median, const msg_series_levels *lv, void *reduced) {
  const int L = p->precision_level;
  const int N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = sizeof(S) == 8 ? MSG_F64 : MSG_F32;
  const S *tpl = (const S *)templates, *ref = (const S *)reference;
  msg_solve_params sp;
  sp.lambda = p->lambda;
  sp.lambda_factor = p->lambda_factor;
  sp.lambda_ref_level = p->stop_level;
  sp.max_iter = p->max_iter;
  sp.stop_eps = p->stop_eps;
  sp.resample_prolong = p->resample_prolong;
  sp.spec = p->spec;
  const bool reuse = pairwise_in != nullptr;
  const int npw = reuse ? 1 : n;  // pairwise solves to run (pairs 0..npw-1)
  const S *init0 = (const S *)p->pair0_init;
  const bool chain = !p->dont_accumulate;
  // Batches: with a chain to overlap, about four (the chain starts when the first is done), and no
  // more pairs per batch than keep the batch's fields (about 15 fields of N^2 per pair) within the
  // 256 MB last-level cache beside the chain's (measured at 32 x 1024^2: batches of 4 pairs take
  // 1.21 s per stage, of 8 pairs 1.29 s, of 2 pairs 1.24 s); otherwise one.
  int b = p->pairwise_batch;
  if (b <= 0) {
    const double pair_bytes = 15.0 * (double)nn * sizeof(S);
    const int cache_pairs = std::max(2, (int)((256.0 * (1 << 20)) / pair_bytes));
    b = chain ? std::min((npw + 3) / 4, cache_pairs) : npw;
  }
  b = std::max(1, std::min(b, npw));
  const bool overlap = chain && b < npw;

  hipStream_t s = nullptr;  // the caller thread's stream for the reconstruction
  {
    std::lock_guard<std::mutex> g(g_series_counts.mu);
    g_series_counts.v[0] = g_series_counts.v[1] = g_series_counts.v[2] = 0;
    g_series_counts.w[0] = g_series_counts.w[1] = 0;
  }
  msg_ctx *rc = nullptr;
  if (chain) CHK(ctx_acquire(L, 1, dt, &rc));
  CtxGuard grc{rc};
  if (chain) CHK(msg_ctx_set_data_term(rc, p->data_term));
  if (chain && p->data_term == MSG_XSHIFT) CHK(msg_ctx_set_y_reg_factor(rc, p->y_reg_factor));

  // Buffers of the stage come from the calling thread's scratch slots: allocated by the first call,
  // grown when a larger series comes, kept for the next call (no allocation in a repeated stage).
  void *pwv = nullptr;
  CHK(scratch_get(100, sizeof(S) * 2 * nn * n, &pwv));
  S *pwp = (S *)pwv;
  for (int i = 0; i < n; ++i) energies[i] = NAN;
  if (reuse && n > 1) {
    HIPCHK(hipMemcpy(pwp + 2 * nn, (const S *)pairwise_in + 2 * nn, sizeof(S) * 2 * nn * (n - 1),
                     hipMemcpyDeviceToDevice));
  }

  // the start_level results of the pairwise solves (altStartLevel replaces some of them); record
  // blocks of the repeated solves come after the pairwise, refine and inverse ones
  const bool alt = p->alt_start_level > 0 && p->alt_start_level != p->start_level;
  DevBuf firstBuf;
  S *firstp = (S *)p->first_try_out;
  if (!firstp && alt) {
    HIPCHK(hipMalloc(&firstBuf.p, sizeof(S) * 2 * nn * n));
    firstp = (S *)firstBuf.p;
  }
  const int nlp = p->stop_level - p->start_level + 1, nlr = p->refine_stop_level - p->refine_start_level + 1;
  const long long nref = chain ? (long long)(n - 1) + (p->calc_inverse ? n : 0) : 0;
  const long long alt_rec0 = (long long)npw * nlp + nref * nlr;
  const size_t alt_field0 = (size_t)npw * level_block(p->start_level, p->stop_level) +
                            (size_t)nref * (chain ? level_block(p->refine_start_level, p->refine_stop_level) : 0);
  if (firstp && reuse && n > 1)  // frames taken from the previous stage were not solved again
    HIPCHK(hipMemcpy(firstp + 2 * nn, (const S *)pairwise_in + 2 * nn, sizeof(S) * 2 * nn * (n - 1),
                     hipMemcpyDeviceToDevice));

  Progress pr;
  if (overlap) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    worker()->post([=, &pr, &sp] {
      const hipError_t e = hipSetDevice(dev);
      if (e != hipSuccess) {
        std::lock_guard<std::mutex> g(pr.mu);
        pr.rc = MSG_ERR_HIP;
        pr.err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        pr.finished = true;
        pr.cv.notify_all();
        return;
      }
      pairwise_all<S>(p, sp, tpl, ref, npw, b, pwp, energies, lv, init0, &pr, firstp, alt_rec0, alt_field0);
    });
  } else {
    pairwise_all<S>(p, sp, tpl, ref, npw, b, pwp, energies, lv, init0, &pr, firstp, alt_rec0, alt_field0);
    if (pr.rc != MSG_OK) return fail(pr.rc, pr.err);
  }
  // Waits (on the calling thread) until pairwise solve i is done; false after a batch failed.
  auto wait_pairwise = [&](int i) {
    std::unique_lock<std::mutex> l(pr.mu);
    pr.cv.wait(l, [&] { return pr.done > i || pr.rc != MSG_OK || pr.finished; });
    return pr.done > i;
  };
  // Stops the batches after an error on the chain and waits for the worker to let go of the buffers.
  auto join = [&](int code) {
    if (overlap) {
      std::unique_lock<std::mutex> l(pr.mu);
      pr.abort = true;
      pr.cv.wait(l, [&] { return pr.finished; });
    }
    return code;
  };

  S *defs = (S *)deformations;
  if (!chain) {
    HIPCHK(hipMemcpy(defs, pwp, sizeof(S) * 2 * nn * n, hipMemcpyDeviceToDevice));
    // with altStartLevel the reference averages frames i > 0 with the start_level results
    // (createDeformationBaseFileName, matchSeries.h:55-62)
    if (alt && n > 1)
      HIPCHK(hipMemcpy(defs + 2 * nn, firstp + 2 * nn, sizeof(S) * 2 * nn * (n - 1), hipMemcpyDeviceToDevice));
  } else {
    // --- sequential chain: accumulate by composition, refine against the reference ---
    hipStream_t rs = rc->stream;
    s = rs;
    const int code = [&]() -> int {
      HIPCHK(hipMemcpyAsync(rc->R[L], ref, nn * sizeof(S), hipMemcpyDeviceToDevice, rs));
      if (!p->dont_normalize) CHK(normalize_impl<S>(rc->R[L], N, 1, p->saturation_pct, p->normalize_min_to_zero, rs));
      if (p->pre_smooth_sigma > 0) CHK(h1_impl<S>(rc->R[L], rc->R[L], N, 1, p->pre_smooth_sigma / (double)(N - 1), rs));
      CHK(rebuild_pyramid<S>(rc, rc->R, -1, 1));
      DevBuf swp, zero;
      struct {
        void *p = nullptr;  // a scratch slot of this thread: not freed here
      } acc, prev;
      CHK(scratch_get(101, sizeof(S) * 2 * nn, &acc.p));
      CHK(scratch_get(102, sizeof(S) * 2 * nn, &prev.p));
      HIPCHK(hipMemsetAsync(acc.p, 0, sizeof(S) * 2 * nn, rs));
      if (p->keep_start_transformation) {
        HIPCHK(hipMalloc(&zero.p, sizeof(S) * 2 * nn));
        HIPCHK(hipMemsetAsync(zero.p, 0, sizeof(S) * 2 * nn, rs));
      }
      if (p->calc_inverse) HIPCHK(hipMalloc(&swp.p, sizeof(S) * nn));
      for (int i = 0; i < n; ++i) {
        if (i < npw && !wait_pairwise(i)) return pr.rc;
        {
          // keep_start_transformation: a pairing solved here hands back the zero it started from, a
          // loaded one (reuseStage1Results) the loaded deformation
          const S *cur = (p->keep_start_transformation && i < npw) ? (const S *)zero.p : pwp + (size_t)i * 2 * nn;
          HIPCHK(hipMemcpyAsync(prev.p, acc.p, sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, rs));
          CHK(compose_impl<S>(cur, prev.p, acc.p, N, rs));
        }
        if (i > 0) {
          HIPCHK(hipMemcpyAsync(rc->T[L], tpl + (size_t)i * nn, nn * sizeof(S), hipMemcpyDeviceToDevice, rs));
          if (!p->dont_normalize)
            CHK(normalize_impl<S>(rc->T[L], N, 1, p->saturation_pct, p->normalize_min_to_zero, rs));
          if (p->pre_smooth_sigma > 0)
            CHK(h1_impl<S>(rc->T[L], rc->T[L], N, 1, p->pre_smooth_sigma / (double)(N - 1), rs));
          CHK(rebuild_pyramid<S>(rc, rc->T, -1, 1));
          HIPCHK(hipMemcpyAsync(rc->phi[L], acc.p, sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, rs));
          CHK(rebuild_pyramid<S>(rc, rc->phi, -1, 2));
          double e = 0;
          LevelOut lo;
          lo.lv = lv;
          lo.rec0 = (long long)npw * (p->stop_level - p->start_level + 1) +
                    (long long)(i - 1) * (p->refine_stop_level - p->refine_start_level + 1);
          lo.field0 = (size_t)npw * level_block(p->start_level, p->stop_level) +
                      (size_t)(i - 1) * level_block(p->refine_start_level, p->refine_stop_level);
          CHK(solve_impl<S>(rc, p->refine_start_level, p->refine_stop_level, &sp, &e, &lo));
          energies[i] = e;
          // keep_start_transformation: the refinement's start stays the accumulated deformation; its
          // result only goes to the frame's average
          HIPCHK(hipMemcpyAsync(p->keep_start_transformation ? (void *)(defs + (size_t)i * 2 * nn) : acc.p,
                                rc->phi[L], sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, rs));
        }
        if (p->calc_inverse) {
          // calcInverseDeformation (matchSeries.h:247-259): reference and template swapped, the
          // solve over the refine levels started from the approximate inverse of the accumulated
          // deformation. The swap copies the pyramids' contents (captured level graphs hold the
          // buffers' addresses) and is undone afterwards.
          if (i == 0) {
            HIPCHK(hipMemcpyAsync(rc->T[L], tpl, nn * sizeof(S), hipMemcpyDeviceToDevice, rs));
            if (!p->dont_normalize)
              CHK(normalize_impl<S>(rc->T[L], N, 1, p->saturation_pct, p->normalize_min_to_zero, rs));
            if (p->pre_smooth_sigma > 0)
              CHK(h1_impl<S>(rc->T[L], rc->T[L], N, 1, p->pre_smooth_sigma / (double)(N - 1), rs));
            CHK(rebuild_pyramid<S>(rc, rc->T, -1, 1));
          }
          auto swap_images = [&]() -> int {
            for (int l = 0; l <= L; ++l) {
              const size_t bytes = sizeof(S) * ((size_t)1 << (2 * l));
              HIPCHK(hipMemcpyAsync(swp.p, rc->R[l], bytes, hipMemcpyDeviceToDevice, rs));
              HIPCHK(hipMemcpyAsync(rc->R[l], rc->T[l], bytes, hipMemcpyDeviceToDevice, rs));
              HIPCHK(hipMemcpyAsync(rc->T[l], swp.p, bytes, hipMemcpyDeviceToDevice, rs));
            }
            return MSG_OK;
          };
          CHK(swap_images());
          CHK(inverse_impl<S>((const S *)acc.p, (S *)rc->phi[L], N, rs));
          CHK(rebuild_pyramid<S>(rc, rc->phi, -1, 2));
          const long long nlr = p->refine_stop_level - p->refine_start_level + 1;
          LevelOut lo;
          lo.lv = lv;
          lo.rec0 = (long long)npw * (p->stop_level - p->start_level + 1) + (long long)(n - 1 + i) * nlr;
          lo.field0 = (size_t)npw * level_block(p->start_level, p->stop_level) +
                      (size_t)(n - 1 + i) * level_block(p->refine_start_level, p->refine_stop_level);
          double e = 0;
          CHK(solve_impl<S>(rc, p->refine_start_level, p->refine_stop_level, &sp, &e, &lo));
          if (p->inverse_out)
            HIPCHK(hipMemcpyAsync((S *)p->inverse_out + (size_t)i * 2 * nn, rc->phi[L], sizeof(S) * 2 * nn,
                                  hipMemcpyDeviceToDevice, rs));
          CHK(swap_images());
        }
        if (!(p->keep_start_transformation && i > 0))  // frame i > 0 was stored by its refinement
          HIPCHK(hipMemcpyAsync(defs + (size_t)i * 2 * nn, (i > 0) ? acc.p : (const void *)pwp, sizeof(S) * 2 * nn,
                                hipMemcpyDeviceToDevice, rs));
      }
      HIPCHK(hipStreamSynchronize(rs));
      return MSG_OK;
    }();
    join(code);
    if (code != MSG_OK) return code == pr.rc && !pr.err.empty() ? fail(code, pr.err) : code;
    if (pr.rc != MSG_OK) return fail(pr.rc, pr.err);
  }
  if (pairwise_out) HIPCHK(hipMemcpy(pairwise_out, pwp, sizeof(S) * 2 * nn * n, hipMemcpyDeviceToDevice));

  // --- reconstruction: warp the raw templates (Inf outside), then mean / count / median ---
  struct {
    void *p = nullptr;  // a scratch slot of this thread: not freed here
  } stack;
  CHK(scratch_get(103, sizeof(S) * nn * std::max(n, 1), &stack.p));
  // reduceDeformations (matchSeries.h:565-570, 678-704): the reduced deformation is applied to every
  // template first (DeformImage, +Inf outside), then the template's own deformation (L386-392)
  DevBuf red, pre;
  // the frames as the reconstruction reads them: raw (raw_templates, when the caller prepared the
  // matching inputs itself), else the templates given
  const S *raw = p->raw_templates ? (const S *)p->raw_templates : tpl;
  const S *src = raw;
  if (p->reduce_deformations) {
    void *rp = reduced;
    if (!rp) {
      HIPCHK(hipMalloc(&red.p, sizeof(S) * 2 * nn));
      rp = red.p;
    }
    HIPCHK(hipStreamSynchronize(s));
    CHK(reduce_impl<S>(defs, n, L, p->reduce_max_iter, p->reduce_stop_eps, rp, nullptr, nullptr));
    HIPCHK(hipMalloc(&pre.p, sizeof(S) * nn * n));
    const size_t cnt = nn * n;
    hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, raw, (const S *)rp,
                       (S *)pre.p, N, n, 1, 0, (double)INFINITY, 0);
    HIPCHK(hipGetLastError());
    src = (const S *)pre.p;
  }
  // averageSeries (matchSeries.h:353-451): DeformImage of each template, +Inf (no sample) or the
  // image's own mean outside the domain, optionally nearest-neighbour
  const int nearest = p->average_nearest ? 1 : 0;
  if (p->average_ext_mean) {
    for (int i = 0; i < n; ++i) {
      double ext = 0;
      CHK(mean_value<S>(src + (size_t)i * nn, (int)nn, &ext, s));
      hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, src + (size_t)i * nn,
                         (const S *)defs + (size_t)i * 2 * nn, (S *)stack.p + (size_t)i * nn, N, 1, 0, 0, ext, nearest);
      HIPCHK(hipGetLastError());
    }
  } else {
    const size_t cnt = nn * n;
    hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, src, (const S *)defs,
                       (S *)stack.p, N, n, 0, 0, (double)INFINITY, nearest);
    HIPCHK(hipGetLastError());
  }
  double fill = 0;
  CHK(mean_value<S>(raw + (size_t)(n - 1) * nn, (int)nn, &fill, s));
  CHK(stats_impl<S>(stack.p, n, N, fill, mean, count, median, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSG_OK;
}

#include "ms_batch.h"

}  // namespace

extern "C" {

int msg_compose(int dtype, const void *phi1, const void *phi2, void *result, int n, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!phi1 || !phi2 || !result) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64) CHK(compose_impl<double>(phi1, phi2, result, n, strm(stream)));
  else CHK(compose_impl<float>(phi1, phi2, result, n, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_series_stats(int dtype, const void *stack, int nimg, int n, double fill, void *mean, void *count,
                     void *median, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!stack || !mean || !count || !median || nimg < 0 || nimg > msg::ST_MAXF)
    return fail(MSG_ERR_ARG, "invalid argument (at most 8192 images)");
  if (dtype == MSG_F64) CHK(stats_impl<double>(stack, nimg, n, fill, mean, count, median, strm(stream)));
  else CHK(stats_impl<float>(stack, nimg, n, fill, mean, count, median, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_resample_rect(int dtype, const void *src, int w, int h, void *dst, int n, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!src || !dst || w < 2 || h < 2) return fail(MSG_ERR_ARG, "invalid argument (source of at least 2 x 2)");
  hipStream_t s = strm(stream);
  // ResampleCopier's positions: the steps divided in double, held in long double, products rounded
  std::vector<double> c(2 * (size_t)n);
  const long double hx = (long double)((double)(w - 1) / (double)(n - 1));
  const long double hy = (long double)((double)(h - 1) / (double)(n - 1));
  for (int k = 0; k < n; ++k) {
    c[k] = (double)(k * hx);
    c[n + k] = (double)(k * hy);
  }
  DevBuf dc;
  HIPCHK(hipMalloc(&dc.p, sizeof(double) * 2 * n));
  HIPCHK(hipMemcpyAsync(dc.p, c.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
  const size_t cnt = (size_t)n * n;
  const double *cx = (const double *)dc.p, *cy = cx + n;
  if (dtype == MSG_F64)
    hipLaunchKernelGGL((msg::k_resample_rect<double>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       (const double *)src, w, h, (double *)dst, n, cx, cy);
  else
    hipLaunchKernelGGL((msg::k_resample_rect<float>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       (const float *)src, w, h, (float *)dst, n, cx, cy);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));  // the coordinates are freed on return
  return MSG_OK;
}

int msg_approx_inverse(int dtype, const void *phi, void *inv, int n, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!phi || !inv || phi == inv) return fail(MSG_ERR_ARG, "invalid argument (phi and inv must be distinct)");
  if (dtype == MSG_F64) CHK(inverse_impl<double>((const double *)phi, (double *)inv, n, strm(stream)));
  else CHK(inverse_impl<float>((const float *)phi, (float *)inv, n, strm(stream)));
  return MSG_OK;
}

int msg_register_translation(int dtype, const void *reference, const void *templ, int n, int max_shift, int *shift,
                             double *corr, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!reference || !templ || !shift) return fail(MSG_ERR_ARG, "invalid argument");
  TranslationWork wk;
  if (dtype == MSG_F64)
    CHK(translation_impl<double>((const double *)reference, (const double *)templ, n, max_shift, shift, corr, wk,
                                 strm(stream)));
  else
    CHK(translation_impl<float>((const float *)reference, (const float *)templ, n, max_shift, shift, corr, wk,
                                strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_normalize(int dtype, void *img, int n, int nimg, double pct, int min_to_zero, void *stream) {
  CHK(check_dtype(dtype));
  CHK(check_n(n));
  if (!img || nimg < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (dtype == MSG_F64) CHK(normalize_impl<double>(img, n, nimg, pct, min_to_zero, strm(stream)));
  else CHK(normalize_impl<float>(img, n, nimg, pct, min_to_zero, strm(stream)));
  HIPCHK(hipStreamSynchronize(strm(stream)));
  return MSG_OK;
}

int msg_match_series_batch(int dtype, int njobs, const msg_series_job *jobs) {
  CHK(check_dtype(dtype));
  if (njobs < 1 || !jobs) return fail(MSG_ERR_ARG, "invalid argument");
  for (int k = 0; k < njobs; ++k) {
    const msg_series_job &j = jobs[k];
    if (!j.params || !j.templates || !j.reference || !j.deformations || !j.energies || !j.mean || !j.count ||
        !j.median || j.n < 1)
      return fail(MSG_ERR_ARG, "job " + std::to_string(k) + ": invalid argument");
    CHK(check_batch_params(j.params));
    if (j.n > msg::ST_MAXF) return fail(MSG_ERR_ARG, "job " + std::to_string(k) + ": at most 8192 templates");
  }
  return dtype == MSG_F64 ? match_series_batch_impl<double>(njobs, jobs) : match_series_batch_impl<float>(njobs, jobs);
}

int msg_series_last_counts(long long *iterations, long long *evals, long long *passes, double *node_iterations,
                           double *node_evals) {
  std::lock_guard<std::mutex> g(g_series_counts.mu);
  if (iterations) *iterations = g_series_counts.v[0];
  if (evals) *evals = g_series_counts.v[1];
  if (passes) *passes = g_series_counts.v[2];
  if (node_iterations) *node_iterations = g_series_counts.w[0];
  if (node_evals) *node_evals = g_series_counts.w[1];
  return MSG_OK;
}

int msg_match_series(const msg_series_params *p, int dtype, const void *templates, int n, const void *reference,
                     const void *pairwise_in, void *deformations, void *pairwise_out, double *energies, void *mean,
                     void *count, void *median, const msg_series_levels *levels, void *reduced) {
  if (!p || !templates || !reference || !deformations || !energies || !mean || !count || !median || n < 1)
    return fail(MSG_ERR_ARG, "invalid argument");
  CHK(check_dtype(dtype));
  const int L = p->precision_level;
  if (L < 1 || L > MAXL) return fail(MSG_ERR_ARG, "precision_level out of range");
  if (p->start_level < 1 || p->stop_level > L || p->start_level > p->stop_level)
    return fail(MSG_ERR_ARG, "invalid startLevel/stopLevel");
  if (!p->dont_accumulate && n > 1 &&
      (p->refine_start_level < 1 || p->refine_stop_level > L || p->refine_start_level > p->refine_stop_level))
    return fail(MSG_ERR_ARG, "invalid refineStartLevel/refineStopLevel");
  // checked before any solve: the reconstruction's median holds a pixel's samples in LDS
  if (n > msg::ST_MAXF) return fail(MSG_ERR_ARG, "series longer than 8192 frames are not supported by the median kernel");
  if (dtype == MSG_F64)
    return match_series_impl<double>(p, templates, n, reference, pairwise_in, deformations, pairwise_out, energies,
                                     mean, count, median, levels, reduced);
  return match_series_impl<float>(p, templates, n, reference, pairwise_in, deformations, pairwise_out, energies, mean,
                                  count, median, levels, reduced);
}

int msg_reduce_deformations(int dtype, const void *deformations, int n, int level, int max_iter, double stop_eps,
                            void *phi, double *energy, int *iterations) {
  CHK(check_dtype(dtype));
  if (!deformations || !phi || n < 1) return fail(MSG_ERR_ARG, "invalid argument");
  if (level < 1 || level > MAXL) return fail(MSG_ERR_ARG, "level out of range");
  if (dtype == MSG_F64) return reduce_impl<double>(deformations, n, level, max_iter, stop_eps, phi, energy, iterations);
  return reduce_impl<float>(deformations, n, level, max_iter, stop_eps, phi, energy, iterations);
}

}  // extern "C"

extern "C" {

int msg_profile_enable(unsigned mask) {
  std::lock_guard<std::mutex> g(g_prof.mu);
  prof_resolve();
  g_prof.mask = mask;
  return MSG_OK;
}

int msg_profile_reset(void) {
  std::lock_guard<std::mutex> g(g_prof.mu);
  prof_resolve();
  for (int k = 0; k < KC_COUNT; ++k) {
    g_prof.ms[k] = 0;
    g_prof.launches[k] = 0;
    g_prof.bytes[k] = 0;
  }
  return MSG_OK;
}

int msg_profile_read(int cls, double *ms, long long *launches, double *bytes) {
  if (cls < 0 || cls >= KC_COUNT) return fail(MSG_ERR_ARG, "unknown kernel class");
  std::lock_guard<std::mutex> g(g_prof.mu);
  prof_resolve();
  if (ms) *ms = g_prof.ms[cls];
  if (launches) *launches = g_prof.launches[cls];
  if (bytes) *bytes = g_prof.bytes[cls];
  return MSG_OK;
}

}  // extern "C"

extern "C" {

int msg_release_cache(void) {
  std::lock_guard<std::mutex> g(g_ctx_cache.mu);
  for (msg_ctx *c : g_ctx_cache.free) msg_ctx_destroy(c);
  g_ctx_cache.free.clear();
  return MSG_OK;
}

}  // extern "C"

extern "C" {

int msg_ctx_set_mode(msg_ctx *c, int mode) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (mode < 0 || mode > 2) return fail(MSG_ERR_ARG, "mode must be 0, 1 or 2");
  c->mode = mode;
  return MSG_OK;
}

int msg_ctx_set_data_term(msg_ctx *c, int data_term) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (data_term != MSG_SSD && data_term != MSG_NCC && data_term != MSG_XSHIFT)
    return fail(MSG_ERR_ARG, "data term must be MSG_SSD, MSG_NCC or MSG_XSHIFT");
  if (data_term == MSG_XSHIFT && (1 << c->L) > 2048)
    return fail(MSG_ERR_ARG, "MSG_XSHIFT: rows of at most 2048 nodes (level 11)");
  if (c->data_term != data_term) {
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);  // the captured kernels depend on it
    c->graphs.clear();
  }
  c->data_term = data_term;
  return MSG_OK;
}

int msg_ctx_set_y_reg_factor(msg_ctx *c, double y) {
  if (!c) return fail(MSG_ERR_ARG, "null context");
  if (!(y >= 0) || !std::isfinite(y)) return fail(MSG_ERR_ARG, "yRegFactor must be finite and >= 0");
  if (c->yreg != y) {
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);  // the captured kernels carry it
    c->graphs.clear();
  }
  c->yreg = y;
  return MSG_OK;
}

int msg_set_default_mode(int mode) {
  if (mode < 0 || mode > 2) return fail(MSG_ERR_ARG, "mode must be 0, 1 or 2");
  g_default_mode = mode;
  std::lock_guard<std::mutex> g(g_ctx_cache.mu);
  for (msg_ctx *c : g_ctx_cache.free) c->mode = mode;
  return MSG_OK;
}

int msg_set_default_exact_decisions(int on) {
  if (on != 0 && on != 1) return fail(MSG_ERR_ARG, "on must be 0 or 1");
  g_default_exact = on;
  std::lock_guard<std::mutex> g(g_ctx_cache.mu);
  for (msg_ctx *c : g_ctx_cache.free) {
    if (c->exact_decisions != on)
      for (auto &gr : c->graphs) (void)hipGraphExecDestroy(gr.exec);  // the captured controls carry it
    if (c->exact_decisions != on) c->graphs.clear();
    c->exact_decisions = on;
  }
  return MSG_OK;
}

}  // extern "C"
