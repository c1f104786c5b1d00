// gfx950 kernels of the batched registration level solve. S is the storage type of the fields
// (float or double); arithmetic is double. A "pair" is one (reference, template) registration
// problem; pair p's images start at p*n*n and its fields at p*2*n*n.
#pragma once

#include "ms_device.h"
#include "ms_state.h"

namespace msg {

constexpr double TAU_MIN = 9.313225746154785e-10;  // 2^-30 (ArmijoSearch.h:64-73)
constexpr double TAU_MAX = 33554432.0;             // 2^25
constexpr double ARMIJO_SIGMA = 0.5;

// ================================================================================================
// Control: one workgroup per pair reduces the partial sums of the previous pass (fixed order) and
// advances the pair's state machine. Phases:
// ================================================================================================
// PH_NCC_REF0 / PH_NCC_REF1: the reference's mean, then its centred moments (NCC, once per level).
// An accepted step goes to the step slot of its iteration (iterations alternate between two slots);
// the update kernel of that iteration applies it and the next iteration's PH_ARMIJO_INIT clears the
// slot, so no phase of its own is needed after the update.
// PH_ARMIJO_INIT also forms the Fletcher-Reeves coefficient from the V-cycle's partial sums; the CG
// direction itself is never stored before the step: the energy passes and the update form it on
// the fly from z and the previous direction, and the update stores it for the next iteration.
enum CtrlPhase { PH_LEVEL_INIT = 0, PH_ARMIJO_INIT = 2, PH_ARMIJO_STEP = 3, PH_NCC_REF0 = 5, PH_NCC_REF1 = 6 };

constexpr int CTRL_THREADS = 256;  // threads of a control workgroup (a small workgroup finds a free slot sooner)
constexpr int CTRL_MAXT = 16;       // partial arrays one control sweep can reduce

// Sums of NT (<= 8) consecutive partial arrays base[t * len .. t * len + len) in the control
// workgroup, in a fixed order. The partials were just written by a kernel on other XCDs, so a load is
// a memory round trip: 256 threads issue all their loads (CTRL_RUN per array and sweep) before
// summing, from clamped indices with out-of-range values dropped by a select, so none waits for
// another. Each thread's sums go to LDS; 16 threads per array then sum 16 of them each, and one
// thread per array the 16 results -- plain LDS reads and adds (cross-lane shuffles of doubles by 16
// waves at once cost more than the loads). The sums land in out[0 .. NT) (shared memory), valid for
// every thread after the call; red / red2 are the caller's LDS blocks of RED_THREADS and 16 doubles
// per array.
constexpr int CTRL_RUN = 4;
constexpr int RED_THREADS = 256;
template <int NT>
__device__ __forceinline__ void reduce_arrays_n(const double *base, int len, double *out, double (*red)[RED_THREADS],
                                                double (*red2)[16]) {
  const int tid = threadIdx.x;
  if (tid < RED_THREADS) {
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 0.;
    for (int i0 = 0; i0 < len; i0 += RED_THREADS * CTRL_RUN) {
      double v[NT][CTRL_RUN];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < CTRL_RUN; ++r) {
          const int k = i0 + r * RED_THREADS + tid;
          v[t][r] = base[(size_t)t * len + (k < len ? k : len - 1)];
        }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < CTRL_RUN; ++r) {
          const int k = i0 + r * RED_THREADS + tid;
          acc[t] += (k < len) ? v[t][r] : 0.;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) red[t][tid] = acc[t];
  }
  __syncthreads();
  if (tid < NT * 16) {
    const int t = tid >> 4, g = tid & 15;
    double a = 0.;
#pragma unroll
    for (int j = 0; j < 16; ++j) a += red[t][16 * g + j];
    red2[t][g] = a;
  }
  __syncthreads();
  if (tid < NT) {
    double a = 0.;
#pragma unroll
    for (int j = 0; j < 16; ++j) a += red2[tid][j];
    out[tid] = a;
  }
  __syncthreads();
}

// nt (<= CTRL_MAXT) arrays: sweeps of at most 8 arrays
__device__ __forceinline__ void reduce_arrays(const double *base, int nt, int len, double *out) {
  __shared__ double red[8][RED_THREADS];
  __shared__ double red2[8][16];
  static_assert(CTRL_THREADS >= RED_THREADS, "the reduction's threads are part of the control workgroup");
  for (int t0 = 0; t0 < nt; t0 += 8) {
    const int m = nt - t0 < 8 ? nt - t0 : 8;
    const double *b = base + (size_t)t0 * len;
    switch (m) {
      case 8: reduce_arrays_n<8>(b, len, out + t0, red, red2); break;
      case 7: reduce_arrays_n<7>(b, len, out + t0, red, red2); break;
      case 6: reduce_arrays_n<6>(b, len, out + t0, red, red2); break;
      case 5: reduce_arrays_n<5>(b, len, out + t0, red, red2); break;
      case 4: reduce_arrays_n<4>(b, len, out + t0, red, red2); break;
      case 3: reduce_arrays_n<3>(b, len, out + t0, red, red2); break;
      case 2: reduce_arrays_n<2>(b, len, out + t0, red, red2); break;
      default: reduce_arrays_n<1>(b, len, out + t0, red, red2); break;
    }
  }
}

struct CtrlArgs {
  int phase;
  int nblk_energy;   // partial count per candidate and term
  int nblk_norm;     // partials per field and term of the V-cycle's last pass
  double lambda, stop_eps;
  int max_iter;
  int spec;          // candidates per energy pass (1..MAX_CAND)
  int trace_cap;
  int data_term;     // DT_SSD, DT_NCC or DT_CONS
  int npairs;        // pairs of the batch (NCC: the reference statistics follow the candidates' partials)
  int sigma_fixed;   // 1: filter width fixed at sigma0 (H1GradientDescent, gradientDescent.h:639-650)
  double sigma0;
  int grad_stop;     // 1: USE_GRADIENT_BASED_STOPPING (gradientDescent.h:416)
  int slot;          // step slot of the current iteration (0 or 1)
  double exact_tol;  // > 0 (FP32 SSD passes): relative error bound of an evaluation; decisions whose
                     // margin lies within it are taken on FP64-arithmetic values (0: never)
};
// Partial sums of the V-cycle's last pass, per field and block: <z,z>, <DE,z>, <DE,d_old>, <DE,DE>.
constexpr int NORM_TERMS = 4;

// The search direction at element i: the smoothed gradient z, or the Fletcher-Reeves direction
// d = z + beta d_old (calcTaudAndDescentDir, gradientDescent.h:218-226) as the FP64 sum rounded to S,
// with d_old taken as zero after a fall-back to the gradient.
template <typename S>
__device__ __forceinline__ S search_dir(const S *__restrict__ z, const S *__restrict__ dold, size_t i, bool cg,
                                        bool use_old, double beta) {
  const S zv = z[i];
  if (!cg) return zv;
  const double od = use_old ? (double)dold[i] : 0.0;
  return static_cast<S>((double)zv + beta * od);
}
// The same from values already loaded (z_i, d_old_i); d_old_i is ignored unless use_old.
template <typename S>
__device__ __forceinline__ S search_dir_val(S zv, S ov, bool cg, bool use_old, double beta) {
  if (!cg) return zv;
  const double od = use_old ? (double)ov : 0.0;
  return static_cast<S>((double)zv + beta * od);
}
template <>
__device__ __forceinline__ float search_dir_val<float>(float zv, float ov, bool cg, bool use_old, double beta) {
  const float dv = __fmaf_rn((float)beta, ov, zv);
  return (cg && use_old) ? dv : zv;
}
// FP32 fields: one fused multiply-add in FP32 (the same expression in every kernel that forms d);
// both loads unconditional and the choice a select, so the loads of neighbouring nodes overlap.
template <>
__device__ __forceinline__ float search_dir<float>(const float *__restrict__ z, const float *__restrict__ dold,
                                                   size_t i, bool cg, bool use_old, double beta) {
  const float zv = z[i], ov = dold[i];
  const float dv = __fmaf_rn((float)beta, ov, zv);
  return (cg && use_old) ? dv : zv;
}

// Where the energy pass stores its partial sums. SSD: 2 terms per candidate (E_SSD, E_reg). NCC: 4
// per candidate (1'Mw, w'Mw, w'Mu, E_reg with w = t - c, u = r - m_r) and, after all candidates of
// all pairs, 2 per pair for the reference's statistics.
__device__ __host__ __forceinline__ size_t part_index(int data_term, int pair, int k, int term, int nblk) {
  return ((size_t)(pair * MAX_CAND + k) * (data_term == DT_NCC ? 4 : 2) + term) * nblk;
}
__device__ __host__ __forceinline__ size_t part_index_ref(int npairs, int pair, int term, int nblk) {
  return ((size_t)npairs * MAX_CAND * 4 + (size_t)pair * 2 + term) * nblk;
}
// NCC energy from the moments of one evaluation (NormalizedCrossCorrelationEnergy::computeEnergy,
// registration.h:1953-1964, with normalizeImageForNCC, L1934-1947, written in moments). The template
// t enters shifted, w = t - c, and the reference centred, u = r - m_r, so that no moment cancels:
// s1 = 1'Mw = m_t - c, s2 = w'Mw, s3 = w'Mu; then m_t = c + s1, v_t = sqrt(s2 - s1^2) (1'M1 = 1) and
// (t - m_t)'M(r - m_r) = s3 - s1 * rb with rb = 1'Mu. A deviation within 1e-15 of zero is not divided
// by (appeqAbsolute, aol.h:779-794).
__device__ __host__ __forceinline__ double ncc_value(double s1, double s2, double s3, double rb, double rv,
                                                    double *vt_out) {
  const double vt = sqrt(fmax(s2 - s1 * s1, 0.0));
  *vt_out = vt;
  const double dt = fabs(vt - 0.0) < 1e-15 ? 1.0 : vt;
  const double dr = fabs(rv - 0.0) < 1e-15 ? 1.0 : rv;
  return -((s3 - s1 * rb) / (dt * dr));
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Plan the next energy pass of the Armijo machine: evaluations the sequential algorithm may ask for next
// (at most PLAN_CAND: the first step size with its double and half, or the next three doublings or
// halvings).
constexpr int PLAN_CAND = 3;
static_assert(PLAN_CAND <= MAX_CAND, "a plan fits the pair state's candidate slots");
__device__ __forceinline__ void armijo_plan(PairState &s, int spec) {
  double c[PLAN_CAND];
  int k = 0;
  const double t = s.a_tau;
  if (s.a_phase == AP_FIRST) {
    c[k++] = t;
    c[k++] = 2 * t;
    c[k++] = t > TAU_MIN ? 0.5 * t : t;
  } else if (s.a_phase == AP_GROW) {
    c[k++] = 2 * t;
    c[k++] = 4 * t;
    c[k++] = 8 * t;
  } else {
    double u = t > TAU_MIN ? 0.5 * t : t;
    for (int j = 0; j < 3; ++j) {
      c[k++] = u;
      u = u > TAU_MIN ? 0.5 * u : u;
    }
  }
  if (k > spec) k = spec;
  for (int j = 0; j < k; ++j) s.cand[j] = c[j];
  s.ncand = k;
  s.searching = 1;
  s.passes++;
}

// One energy evaluation: total E, unweighted E_reg and (NCC) the data energy and the warped
// template's mean / deviation.
struct EvalRec {
  double E, R, D, M, V;
  int ex;  // from an FP64-arithmetic pass
};

__device__ __forceinline__ bool armijo_lookup(const PairState &s, double t, EvalRec &ev) {
  for (int j = 0; j < s.ncand; ++j)
    if (s.cand[j] == t) {
      ev = {s.cand_E[j], s.cand_R[j], s.cand_D[j], s.cand_M[j], s.cand_V[j], s.cand_ex};
      return true;
    }
  for (int j = 0; j < 2; ++j)
    if (s.last_tau[j] == t) {
      ev = {s.last_E[j], s.last_R[j], s.last_D[j], s.last_M[j], s.last_V[j], s.last_ex[j]};
      return true;
    }
  return false;
}

__device__ __forceinline__ void armijo_consume(PairState &s, double t, const EvalRec &ev) {
  s.last_tau[1] = s.last_tau[0];
  s.last_E[1] = s.last_E[0];
  s.last_R[1] = s.last_R[0];
  s.last_D[1] = s.last_D[0];
  s.last_M[1] = s.last_M[0];
  s.last_V[1] = s.last_V[0];
  s.last_ex[1] = s.last_ex[0];
  s.last_tau[0] = t;
  s.last_E[0] = ev.E;
  s.last_R[0] = ev.R;
  s.last_D[0] = ev.D;
  s.last_M[0] = ev.M;
  s.last_V[0] = ev.V;
  s.last_ex[0] = ev.ex;
  s.evals++;
}

// A step decision compares E = E(phi + t d) with f = E(phi): Armijo's G >= 1/2, i.e.
// E - f <= 1/2 Df t, and f >= E (ArmijoSearch.h:148-190). With FP32 energy passes (tol > 0) a margin
// within tol (|f| + |E|) of either test could go either way on rounding alone; then both values are
// evaluated again in FP64 arithmetic at the stored positions before deciding.
__device__ __forceinline__ bool needs_exact(const PairState &s, double f, double t, const EvalRec &ev, double tol) {
  if (tol <= 0 || (ev.ex && s.f_ex) || !isfinite(ev.E) || !isfinite(f)) return false;
  const double b = tol * (fabs(f) + fabs(ev.E));
  return fabs(f - ev.E) <= b || fabs((ev.E - f) - ARMIJO_SIGMA * s.a_Df * t) <= b;
}
// Next pass: FP64 arithmetic at step t and, unless f already comes from such a pass at the current
// position (f_ex: consecutive near ties late in a level solve), at 0 (the current position: f).
__device__ __forceinline__ void plan_exact(PairState &s, double t) {
  s.cand[0] = t;
  s.cand[1] = 0.0;
  s.ncand = s.f_ex ? 1 : 2;
  s.exact = 1;
  s.searching = 1;
  s.passes++;
}

__device__ __forceinline__ void armijo_start(PairState &s, int dirmode, double Df, double oldTau) {
  s.dirmode = dirmode;
  s.a_Df = Df;
  s.a_tau = clampd(oldTau, TAU_MIN, TAU_MAX);
  s.a_phase = AP_FIRST;
  s.ncand = 0;
  s.last_tau[0] = s.last_tau[1] = -1.0;
}

// End of an iteration (GradientDescentBase::applySingle L361-439).
__device__ __forceinline__ void finish_iteration(PairState &s, double t, IterTrace *trace, int trace_cap, int pair, int slot) {
  s.step_tau[slot] = t;
  s.step_dir[slot] = s.dirmode;
  double En = s.energy, Rn = s.ereg_new;
  if (t != 0) {
    EvalRec ev;
    if (armijo_lookup(s, t, ev)) {
      En = ev.E;
      s.f_ex = ev.ex;
      Rn = ev.R;
      s.ncc_d = ev.D;  // the NCC variation at the new position needs these
      s.ncc_m = ev.M;
      s.ncc_v = ev.V;
      if (ev.M == ev.M) s.ncc_c = ev.M;  // later moments about the current mean
    } else {
      En = __builtin_nan("");  // cannot happen: the accepted step is one of the last two evaluations
    }
  }
  const double energy = s.energy;
  s.energy_new = En;
  s.ereg_new = Rn;
  s.tau = t;
  const double sigma_used = s.sigma;
  // the norm is of DE at the position the iteration started from (gradientDescent.h:416)
  const double crit = s.grad_stop ? sqrt(s.dd) : energy - En;
  bool stopping = (crit > s.stop_eps) && (s.iter < s.max_iter);
  if (s.sigma_fixed && t == 0) stopping = false;  // L418-422
  if (!s.sigma_fixed && (0.5 * s.sigma > 0.0001) && ((energy - En) < 0.00001)) {
    s.sigma *= 0.5;
    s.tau = 1.;
    stopping = (s.sigma > 0.0001) && (s.iter < s.max_iter);
  }
  if (trace && s.iter - 1 < trace_cap) {
    IterTrace &tr = trace[(size_t)pair * trace_cap + (s.iter - 1)];
    tr.energy = En;
    tr.ereg = Rn;
    tr.tau = t;
    tr.tau_first = s.tau_first;
    tr.sigma = sigma_used;
    tr.evals = s.evals;
    tr.passes = s.passes;
    tr.beta = s.beta;
  }
  s.total_evals += s.evals;
  s.total_passes += s.passes;
  s.energy = En;
  s.active = stopping ? 1 : 0;
  s.searching = 0;
  s.ncand = 0;
}

// Replays getTimestepWidthWithArmijoLineSearch (ArmijoSearch.h:118-200) over the evaluations
// available, including the fall-back to the smoothed gradient (gradientDescent.h:231-235).
__device__ __forceinline__ void armijo_run(PairState &s, int spec, double tol, IterTrace *trace, int trace_cap, int pair, int slot) {
  const double f = s.energy;
  for (int guard = 0; guard < 256; ++guard) {
    double t = s.a_tau;
    EvalRec ev;
    double result = -1;  // >= 0 when the search ends
    if (s.a_phase == AP_FIRST) {
      if (s.a_Df >= 0 || !isfinite(f)) {
        s.evals++;  // the reference evaluates E(tau0) before these checks
        result = 0;
      } else {
        if (!armijo_lookup(s, t, ev)) {
          armijo_plan(s, spec);
          return;
        }
        if (needs_exact(s, f, t, ev, tol)) {
          plan_exact(s, t);
          return;
        }
        armijo_consume(s, t, ev);
        const double E = ev.E;
        const double G = (E - f) / (s.a_Df * t);
        s.a_phase = (G >= ARMIJO_SIGMA && f >= E) ? AP_GROW : AP_SHRINK;
        continue;
      }
    } else if (s.a_phase == AP_GROW) {
      const double tn = t * 2.;
      if (!armijo_lookup(s, tn, ev)) {
        armijo_plan(s, spec);
        return;
      }
      if (needs_exact(s, f, tn, ev, tol)) {
        plan_exact(s, tn);
        return;
      }
      armijo_consume(s, tn, ev);
      const double E = ev.E;
      s.a_tau = tn;
      const double G = (E - f) / (s.a_Df * tn);
      if (G >= ARMIJO_SIGMA && f >= E && tn <= TAU_MAX) continue;
      s.a_tau *= 0.5;
      result = s.a_tau > TAU_MIN ? s.a_tau : 0.;
    } else {  // AP_SHRINK
      const double tn = t > TAU_MIN ? t * 0.5 : t;
      if (!armijo_lookup(s, tn, ev)) {
        armijo_plan(s, spec);
        return;
      }
      if (needs_exact(s, f, tn, ev, tol)) {
        plan_exact(s, tn);
        return;
      }
      armijo_consume(s, tn, ev);
      const double E = ev.E;
      s.a_tau = tn;
      const double G = (E - f) / (s.a_Df * tn);
      if ((!isfinite(E) || (G < ARMIJO_SIGMA || f < E)) && (tn > TAU_MIN)) continue;
      result = tn > TAU_MIN ? tn : 0.;
    }
    // the search ended with `result`
    if (s.dirmode == DIR_CG) {
      s.tau_first = result;
      if (result == 0) {  // not a descent step along the CG direction: retry along the gradient
        s.olddir_zero = 1;
        armijo_start(s, DIR_GRAD, s.Dfz, 0.0);
        s.ncand = 0;
        continue;
      }
    }
    finish_iteration(s, result, trace, trace_cap, pair, slot);
    return;
  }
  finish_iteration(s, 0.0, trace, trace_cap, pair, slot);  // unreachable guard
}

// Runs one control phase for one pair with the whole workgroup (the reductions need every thread)
// on the workgroup's copy s of the pair state.
__device__ __forceinline__ void control_pair(int pair, PairState &s, const double *__restrict__ part_e,
                             const double *__restrict__ part_n, const CtrlArgs &a, int *__restrict__ flags,
                             IterTrace *__restrict__ trace) {
  __shared__ double sums[CTRL_MAXT];
  const bool participates = s.active && !s.searching;

  if (a.phase == PH_NCC_REF0 || a.phase == PH_NCC_REF1) {
    const int nb = a.nblk_energy;
    reduce_arrays(part_e + part_index_ref(a.npairs, pair, 0, nb), a.phase == PH_NCC_REF1 ? 2 : 1, nb, sums);
    const double r1 = sums[0], r2 = a.phase == PH_NCC_REF1 ? sums[1] : 0;
    if (threadIdx.x == 0) {
      if (a.phase == PH_NCC_REF0) {
        s.ncc_rm = r1;  // 1'Mr
      } else {
        s.ncc_rb = r1;  // 1'M(r - m_r)
        s.ncc_rv = sqrt(fmax(r2, 0.0));
        s.ncc_c = s.ncc_rm;  // the first template moments are taken about the reference's mean
      }
    }
  } else if ((a.phase == PH_LEVEL_INIT || a.phase == PH_ARMIJO_STEP) && a.data_term == DT_NCC) {
    const int nc = s.ncand;
    const int nb = a.nblk_energy;
    reduce_arrays(part_e + part_index(DT_NCC, pair, 0, 0, nb), 4 * nc, nb, sums);  // candidate k, term t at 4k+t
    if (threadIdx.x == 0) {
      for (int k = 0; k < nc; ++k) {
        double vt;
        const double D = ncc_value(sums[4 * k], sums[4 * k + 1], sums[4 * k + 2], s.ncc_rb, s.ncc_rv, &vt);
        double E = 0;
        E += D;
        E += s.lambda * sums[4 * k + 3];  // LinCombOp: (0 + E_NCC) + lambda E_reg
        s.cand_E[k] = E;
        s.cand_R[k] = sums[4 * k + 3];
        s.cand_D[k] = D;
        s.cand_M[k] = s.ncc_c + sums[4 * k];
        s.cand_V[k] = vt;
      }
    }
  } else if ((a.phase == PH_LEVEL_INIT || a.phase == PH_ARMIJO_STEP) && a.data_term == DT_XSHIFT) {
    // rows' NCC energies summed, then divided by the number of rows (FELineShiftRegisEnergy::apply)
    const int nc = s.ncand;
    reduce_arrays(part_e + part_index(DT_SSD, pair, 0, 0, a.nblk_energy), 2 * nc, a.nblk_energy, sums);
    if (threadIdx.x == 0) {
      for (int k = 0; k < nc; ++k) {
        double D = sums[2 * k];
        D /= a.nblk_energy;
        double E = 0;
        E += D;
        E += s.lambda * sums[2 * k + 1];
        s.cand_E[k] = E;
        s.cand_R[k] = sums[2 * k + 1];
        s.cand_D[k] = D;
        s.cand_M[k] = s.cand_V[k] = 0;
      }
      s.cand_ex = 0;
    }
  } else if (a.phase == PH_LEVEL_INIT || a.phase == PH_ARMIJO_STEP) {
    const int nc = s.ncand;
    reduce_arrays(part_e + part_index(DT_SSD, pair, 0, 0, a.nblk_energy), 2 * nc, a.nblk_energy, sums);
    if (threadIdx.x == 0) {
      for (int k = 0; k < nc; ++k) {
        double E = 0;
        E += sums[2 * k];
        E += s.lambda * sums[2 * k + 1];  // LinCombOp: (0 + E_SSD) + lambda E_reg
        s.cand_E[k] = E;
        s.cand_R[k] = sums[2 * k + 1];
        s.cand_D[k] = s.cand_M[k] = s.cand_V[k] = 0;
      }
      s.cand_ex = s.exact;
      if (s.exact) {  // the value at 0 is the FP64-arithmetic energy of the current position
        for (int k = 0; k < nc; ++k)
          if (s.cand[k] == 0.0) s.energy = s.cand_E[k];
        s.f_ex = 1;
        s.exact = 0;
      }
    }
  }
  if (a.phase == PH_LEVEL_INIT || a.phase == PH_ARMIJO_STEP) {
    if (threadIdx.x == 0) {
      if (a.phase == PH_LEVEL_INIT) {
        s.ncc_d = s.cand_D[0];
        s.ncc_m = s.ncc_c = s.cand_M[0];
        s.ncc_v = s.cand_V[0];
        s.energy = s.energy_new = s.energy_start = s.cand_E[0];
        s.ereg_new = s.ereg_start = s.cand_R[0];
        s.tau = 1.0;
        s.sigma = a.sigma_fixed ? a.sigma0 : 1.0;
        s.sigma_fixed = a.sigma_fixed;
        s.grad_stop = a.grad_stop;
        s.dd = 0;
        s.iter = 0;
        s.max_iter = a.max_iter;
        s.stop_eps = a.stop_eps;
        s.active = a.max_iter > 0 ? 1 : 0;
        s.searching = 0;
        s.ncand = 0;
        s.olddir_zero = 1;
        s.zz_old = 0;
        s.total_evals = s.total_passes = 0;
        s.step_tau[0] = s.step_tau[1] = 0;
      } else if (s.searching) {
        armijo_run(s, a.spec, a.exact_tol, trace, a.trace_cap, pair, a.slot);
      }
    }
  } else if (a.phase == PH_ARMIJO_INIT) {
    // the previous iteration's update has run (stream order): free its slot
    if (threadIdx.x == 0) s.step_tau[a.slot ^ 1] = 0;
    if (!participates) return;
    // field f, term t at f * NORM_TERMS + t
    reduce_arrays(part_n + (size_t)(2 * pair) * NORM_TERMS * a.nblk_norm, 2 * NORM_TERMS, a.nblk_norm, sums);
    if (threadIdx.x == 0) {
      double zz = 0, Dfz = 0, Dfd = 0;
      zz += sums[0];
      zz += sums[NORM_TERMS];
      Dfz += sums[1];
      Dfz += sums[NORM_TERMS + 1];
      Dfd += sums[2];
      Dfd += sums[NORM_TERMS + 2];
      s.iter += 1;
      s.beta = ((s.iter - 1) % 10) ? (zz / s.zz_old) : 0.0;
      s.zz_old = zz;
      s.evals = 0;
      s.passes = 0;
      // <DE, d> of d = z + beta d_old, or of d = z when the old direction counts as zero
      s.cg_old = s.olddir_zero ? 0 : 1;
      const double Df = s.cg_old ? Dfz + s.beta * Dfd : Dfz;
      s.Df = Df;
      s.Dfz = Dfz;
      if (a.grad_stop) {
        double dd = 0;
        dd += sums[3];
        dd += sums[NORM_TERMS + 3];
        s.dd = dd;
      }
      s.olddir_zero = 0;  // the update stores this iteration's CG direction
      s.tau_first = 0;
      armijo_start(s, DIR_CG, Df, s.tau);
      s.searching = 1;
      armijo_run(s, a.spec, a.exact_tol, trace, a.trace_cap, pair, a.slot);
    }
  }
  if (threadIdx.x == 0 && flags) flags[pair] = (s.searching ? 1 : 0) | (s.active ? 2 : 0);
}

__global__ void __launch_bounds__(CTRL_THREADS)
k_control(PairState *__restrict__ st, const double *__restrict__ part_e, const double *__restrict__ part_n, CtrlArgs a,
          int *__restrict__ flags, IterTrace *__restrict__ trace) {
  // The state machine runs on an LDS copy of the pair state: one coalesced load and one store
  // instead of a chain of dependent global accesses by the thread that advances it.
  static_assert(sizeof(PairState) % 4 == 0, "pair state copied in 32-bit words");
  constexpr int NW = sizeof(PairState) / 4;
  __shared__ PairState ss;
  const int pair = blockIdx.x;
  unsigned *g = reinterpret_cast<unsigned *>(st + pair), *l = reinterpret_cast<unsigned *>(&ss);
  for (int i = threadIdx.x; i < NW; i += CTRL_THREADS) l[i] = g[i];
  __syncthreads();
  control_pair(pair, ss, part_e, part_n, a, flags, trace);
  __syncthreads();
  for (int i = threadIdx.x; i < NW; i += CTRL_THREADS) g[i] = l[i];
}



// A pair takes part in a new iteration when it is still iterating and no step search is pending
// (device-driven iterations can leave one open across an iteration boundary); the update of its
// last step has run before (stream order).
__device__ __forceinline__ bool participates(const PairState &s) {
  return s.active && !s.searching;
}


// ================================================================================================
// Energy E(phi + tau d) for up to MAX_CAND step sizes per pair, element-centric.
// SSD part: SSDEnergy::evaluateIntegrand (registration.h:1607-1615) integrated as in
// FENonlinIntegrationVectorInterface::applyAdd (FEOpInterface.h:2586-2613).
// Regularisation part: DisplacementLengthEnergy (hyperelastic.h:178-183) = 1/2 sum_c psi_c^T K psi_c,
// assembled element by element. Output: per block partial sums (ssd, reg) per candidate.
// ================================================================================================
constexpr int EN_TX = 64, EN_TY = 4;  // elements per workgroup tile (one per thread)
// The FP32 SSD pass (k_energy_f32, and its FP64 re-evaluation) takes tiles of 64 x 4 * run elements,
// a vertical run of `run` elements per thread: the workgroup's reduction and partial sums are shared by
// the run, and a run reads run + 1 node rows instead of 2 run. Runs of 4 from 1024^2 up; runs of 2
// below, where a single pair would otherwise give too few workgroups to hide the gathers' latency
// (one pair at 512^2: 8.5 us per pass instead of 11.2 us; at 1024^2 runs of 4 are faster, 21.3 us
// against 22.6). The run depends on the grid size only, so a pair's partial sums do not depend on the
// batch it sits in.
constexpr int EN32_RUN_FROM_1024 = 4, EN32_RUN_BELOW = 2;
__host__ __device__ constexpr int en32_run(int n) { return n >= 1024 ? EN32_RUN_FROM_1024 : EN32_RUN_BELOW; }

// The terms of element (ex, ey) of one pair for each candidate step size, added to accS (SSD) and
// accR (regulariser): FP64 arithmetic at the stored positions (step_at). Used by the FP64 pass and by
// the FP64-arithmetic re-evaluation of the FP32 pass (S = float).
template <typename S>
__device__ __forceinline__ void energy_elem(const S *__restrict__ Rp, const S *__restrict__ Tp,
                                            const S *__restrict__ px, const S *__restrict__ py,
                                            const S *__restrict__ zx, const S *__restrict__ zy,
                                            const S *__restrict__ ox_, const S *__restrict__ oy_, int n, double h,
                                            int width, const PairState &sp, int ex, int ey, double (&accS)[MAX_CAND],
                                            double (&accR)[MAX_CAND]) {
  const int nc = sp.ncand;
  const int dm = sp.dirmode;
  const bool cg = (dm == DIR_CG), use_old = sp.cg_old != 0;
  const double beta = sp.beta;
  const bool use_dir = (dm != DIR_NONE);
  double f0[4], f1[4], d0[4], d1[4], rq[4];
  double rn[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int g = l2g(ex, ey, b, n);
    f0[b] = ld(px, g);
    f1[b] = ld(py, g);
    d0[b] = use_dir ? (double)search_dir<S>(zx, ox_, g, cg, use_old, beta) : 0.0;
    d1[b] = use_dir ? (double)search_dir<S>(zy, oy_, g, cg, use_old, beta) : 0.0;
    rn[b] = ld(Rp, g);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double w = 0.;
    for (int b = 0; b < 4; ++b) w += rn[b] * bq(b, q);
    rq[q] = w;
  }
  const double vol = h * h;
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) {
    if (k < nc) {
      const double tau = sp.cand[k];
      double p0[4], p1[4];
      for (int b = 0; b < 4; ++b) {
        p0[b] = use_dir ? (double)step_at<S>((S)f0[b], tau, (S)d0[b]) : f0[b];
        p1[b] = use_dir ? (double)step_at<S>((S)f1[b], tau, (S)d1[b]) : f1[b];
      }
      double a = 0.;
      for (int q = 0; q < 4; ++q) {
        double ox = 0., oy = 0.;
        for (int b = 0; b < 4; ++b) {
          ox += p0[b] * bq(b, q);
          oy += p1[b] * bq(b, q);
        }
        const Hit hp = deform_qp(ex, ey, q, ox, oy, n, h, width);
        double t4[4];
        const double t = eval_img(Tp, hp, n, t4) - rq[q];
        a += 0.5 * (t * t) * 0.25;
      }
      a *= vol;
      accS[k] += a;
      accR[k] += 0.5 * (stiff_form(p0) + stiff_form(p1));
    }
  }
}

// Sums of accS / accR over the workgroup written as the partials of block `tile` (nblk per
// candidate and term) of one pair.
__device__ __forceinline__ void energy_write_partials(double (&accS)[MAX_CAND], double (&accR)[MAX_CAND], int nc,
                                                      double *__restrict__ part, int pair, int tile, int nblk) {
  __shared__ double red[4];
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) {  // static indices keep acc* in registers; nc is uniform
    if (k < nc) {
      const double s = block_sum<256>(accS[k], red);
      const double r = block_sum<256>(accR[k], red);
      if (threadIdx.x == 0) {
        part[((size_t)(pair * MAX_CAND + k) * 2 + 0) * nblk + tile] = s;
        part[((size_t)(pair * MAX_CAND + k) * 2 + 1) * nblk + tile] = r;
      }
    }
  }
}

// One workgroup's tile of the FP64 pass for one pair: one element per thread.
template <typename S>
__device__ __forceinline__ void energy_tile(const S *__restrict__ R, const S *__restrict__ T,
                                            const S *__restrict__ phi, const S *__restrict__ z,
                                            const S *__restrict__ dold, int n, double h, int width,
                                            const PairState *st, double *__restrict__ part, int ntx, int nty,
                                            int pair, int tile) {
  const size_t nn = (size_t)n * n;
  const S *Rp = R + pair * nn, *Tp = T + pair * nn;
  const S *px = phi + pair * 2 * nn, *py = px + nn;
  const S *zx = z + pair * 2 * nn, *zy = zx + nn, *ox_ = dold + pair * 2 * nn, *oy_ = ox_ + nn;
  const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
  const int ey = (tile / ntx) * EN_TY + (threadIdx.x / EN_TX);
  double accS[MAX_CAND], accR[MAX_CAND];
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) accS[k] = accR[k] = 0.0;
  if ((ex < n - 1) && (ey < n - 1))
    energy_elem<S>(Rp, Tp, px, py, zx, zy, ox_, oy_, n, h, width, st[pair], ex, ey, accS, accR);
  energy_write_partials(accS, accR, st[pair].ncand, part, pair, tile, ntx * nty);
}

template <typename S>
__global__ void __launch_bounds__(256)
k_energy(const S *__restrict__ R, const S *__restrict__ T, const S *__restrict__ phi, const S *__restrict__ z,
         const S *__restrict__ dold, int n, double h, int width, const PairState *st,
         double *__restrict__ part, int ntx, int nty, int npairs) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  if (st[pair].ncand == 0) return;  // uniform across the workgroup
  energy_tile<S>(R, T, phi, z, dold, n, h, width, st, part, ntx, nty, pair, tile);
}

// ================================================================================================
// Consistency energy pass of reduceDeformations (matchSeries.h:678-704): for up to MAX_CAND step
// sizes, the sum over the n deformations psi_i (2 fields each) of ConsistencyEnergyOp
// (AmbrosioTortorelli.h:687-735), 1/2 int |psi_i(x + phi(x)) + phi(x)|^2 at the clipped deformed quad
// point. Partials as the SSD pass (term 1, the regulariser, is absent: 0).
// ================================================================================================
template <typename S>
__global__ void __launch_bounds__(256)
k_energy_cons(const S *__restrict__ psi, int npsi, const S *__restrict__ phi, const S *__restrict__ z,
              const S *__restrict__ dold, int n, double h, int width, const PairState *st, double *__restrict__ part,
              int ntx, int nty, int npairs) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  const int nc = st[pair].ncand;
  if (nc == 0) return;  // uniform across the workgroup
  const int dm = st[pair].dirmode;
  __shared__ double red[4];

  const size_t nn = (size_t)n * n;
  const S *px = phi + pair * 2 * nn, *py = px + nn;
  const S *zx = z + pair * 2 * nn, *zy = zx + nn, *ox_ = dold + pair * 2 * nn, *oy_ = ox_ + nn;
  const bool cg = (dm == DIR_CG), use_old = st[pair].cg_old != 0;
  const double beta = st[pair].beta;
  const bool use_dir = (dm != DIR_NONE);

  const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
  const int ey = (tile / ntx) * EN_TY + (threadIdx.x / EN_TX);
  const bool valid = (ex < n - 1) && (ey < n - 1);

  double acc[MAX_CAND];
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) acc[k] = 0.0;

  if (valid) {
    double f0[4], f1[4], d0[4], d1[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int g = l2g(ex, ey, b, n);
      f0[b] = ld(px, g);
      f1[b] = ld(py, g);
      d0[b] = use_dir ? (double)search_dir<S>(zx, ox_, g, cg, use_old, beta) : 0.0;
      d1[b] = use_dir ? (double)search_dir<S>(zy, oy_, g, cg, use_old, beta) : 0.0;
    }
    const double vol = h * h;
#pragma unroll
    for (int k = 0; k < MAX_CAND; ++k) {
      if (k < nc) {
        const double tau = st[pair].cand[k];
        double p0[4], p1[4];
        for (int b = 0; b < 4; ++b) {
          p0[b] = use_dir ? (double)step_at<S>((S)f0[b], tau, (S)d0[b]) : f0[b];
          p1[b] = use_dir ? (double)step_at<S>((S)f1[b], tau, (S)d1[b]) : f1[b];
        }
        double a = 0.;
        for (int q = 0; q < 4; ++q) {
          double ox = 0., oy = 0.;
          for (int b = 0; b < 4; ++b) {
            ox += p0[b] * bq(b, q);
            oy += p1[b] * bq(b, q);
          }
          const Hit hp = deform_qp(ex, ey, q, ox, oy, n, h, width);
          double e = 0.;
          for (int i = 0; i < npsi; ++i) {
            const S *pi = psi + (size_t)i * 2 * nn;
            double u4[4];
            const double vx = eval_img(pi, hp, n, u4) + ox, vy = eval_img(pi + nn, hp, n, u4) + oy;
            e += vx * vx + vy * vy;
          }
          a += 0.5 * e * 0.25;
        }
        acc[k] = a * vol;
      }
    }
  }
  const int nblk = tiles;
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) {
    if (k < nc) {
      const double s = block_sum<256>(acc[k], red);
      if (threadIdx.x == 0) {
        part[((size_t)(pair * MAX_CAND + k) * 2 + 0) * nblk + tile] = s;
        part[((size_t)(pair * MAX_CAND + k) * 2 + 1) * nblk + tile] = 0.0;
      }
    }
  }
}

// ================================================================================================
// NCC energy pass (deformationModel 0) for up to MAX_CAND step sizes per pair, element-centric.
// NormalizedCrossCorrelationEnergy::computeEnergy (registration.h:1953-1964) in moments of the
// assembled mass matrix: with t the nodal FastDeformImage of T at phi + tau d, w = t - c for a shift
// c near the M-mean of t (the last accepted one) and u = r - m_r the centred reference, every element
// adds its share of 1'Mw, w'Mw, w'Mu (the element matrix of the assembled M). Moments about a nearby
// shift keep the differences between the candidates of one pass resolvable in FP64. The control
// combines them (ncc_value).
// ================================================================================================
template <typename S>
__global__ void __launch_bounds__(256)
k_energy_ncc(const S *__restrict__ R, const S *__restrict__ T, const S *__restrict__ phi, const S *__restrict__ z,
             const S *__restrict__ dold, int n, double h, const PairState *st, double *__restrict__ part, int ntx,
             int nty, int npairs) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  const int nc = st[pair].ncand;
  if (nc == 0) return;  // uniform across the workgroup
  const int dm = st[pair].dirmode;
  __shared__ double red[4];

  const size_t nn = (size_t)n * n;
  const S *Rp = R + pair * nn, *Tp = T + pair * nn;
  const S *px = phi + pair * 2 * nn, *py = px + nn;
  const S *zx = z + pair * 2 * nn, *zy = zx + nn, *ox_ = dold + pair * 2 * nn, *oy_ = ox_ + nn;
  const bool cg = (dm == DIR_CG), use_old = st[pair].cg_old != 0;
  const double beta = st[pair].beta;
  const bool use_dir = (dm != DIR_NONE);

  const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
  const int ey = (tile / ntx) * EN_TY + (threadIdx.x / EN_TX);
  const bool valid = (ex < n - 1) && (ey < n - 1);

  double acc[MAX_CAND][4];
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k)
    for (int t = 0; t < 4; ++t) acc[k][t] = 0.0;

  if (valid) {
    const double rm = st[pair].ncc_rm, c = st[pair].ncc_c;
    double f0[4], f1[4], d0[4], d1[4], u[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int g = l2g(ex, ey, b, n);
      f0[b] = ld(px, g);
      f1[b] = ld(py, g);
      d0[b] = use_dir ? (double)search_dir<S>(zx, ox_, g, cg, use_old, beta) : 0.0;
      d1[b] = use_dir ? (double)search_dir<S>(zy, oy_, g, cg, use_old, beta) : 0.0;
      u[b] = ld(Rp, g) + (-rm);  // the reference centred as normalizeImageForNCC does (addToAll)
    }
    double ml[4][4];
    mass_local(h * h, ml);
#pragma unroll
    for (int k = 0; k < MAX_CAND; ++k) {
      if (k < nc) {
        const double tau = st[pair].cand[k];
        double p0[4], p1[4], w[4];
        for (int b = 0; b < 4; ++b) {
          p0[b] = use_dir ? (double)step_at<S>((S)f0[b], tau, (S)d0[b]) : f0[b];
          p1[b] = use_dir ? (double)step_at<S>((S)f1[b], tau, (S)d1[b]) : f1[b];
          w[b] = fast_warp(Tp, n, h, ex + (b & 1), ey + ((b >> 1) & 1), p0[b], p1[b]) - c;
        }
        for (int i = 0; i < 4; ++i) {
          double m = 0.;
          for (int j = 0; j < 4; ++j) m += ml[i][j] * w[j];
          acc[k][0] += m;
          acc[k][1] += w[i] * m;
          acc[k][2] += u[i] * m;
        }
        acc[k][3] = 0.5 * (stiff_form(p0) + stiff_form(p1));
      }
    }
  }
  const int nblk = tiles;
#pragma unroll
  for (int k = 0; k < MAX_CAND; ++k) {  // static indices keep acc in registers; nc is uniform
    if (k < nc) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double v = block_sum<256>(acc[k][t], red);
        if (threadIdx.x == 0) part[part_index(DT_NCC, pair, k, t, nblk) + tile] = v;
      }
    }
  }
}

// The reference's statistics for the NCC pass, once per level: pass 0 the M-mean 1'Mr, pass 1 (with
// the mean in the pair state) the centred moments 1'Mu and u'Mu of u = r - m_r (normalizeImageForNCC,
// registration.h:1934-1947). Partials in the reference slots of the energy layout.
template <typename S>
__global__ void __launch_bounds__(256)
k_ncc_ref(const S *__restrict__ R, int n, double h, const PairState *st, double *__restrict__ part, int ntx, int nty,
          int npairs, int centred) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  __shared__ double red[4];
  const size_t nn = (size_t)n * n;
  const S *Rp = R + pair * nn;
  const int ex = (tile % ntx) * EN_TX + (threadIdx.x % EN_TX);
  const int ey = (tile / ntx) * EN_TY + (threadIdx.x / EN_TX);
  double a1 = 0., a2 = 0.;
  if ((ex < n - 1) && (ey < n - 1)) {
    const double rm = centred ? st[pair].ncc_rm : 0.0;
    double u[4];
    for (int b = 0; b < 4; ++b) u[b] = centred ? ld(Rp, l2g(ex, ey, b, n)) + (-rm) : ld(Rp, l2g(ex, ey, b, n));
    double ml[4][4];
    mass_local(h * h, ml);
    for (int i = 0; i < 4; ++i) {
      double m = 0.;
      for (int j = 0; j < 4; ++j) m += ml[i][j] * u[j];
      a1 += m;
      a2 += u[i] * m;
    }
  }
  const double s1 = block_sum<256>(a1, red);
  const double s2 = block_sum<256>(a2, red);
  if (threadIdx.x == 0) {
    part[part_index_ref(npairs, pair, 0, tiles) + tile] = s1;
    part[part_index_ref(npairs, pair, 1, tiles) + tile] = s2;
  }
}

// ================================================================================================
// Variation DE = F_SSD + lambda K phi, node-centric with the element contributions staged in LDS.
// F_SSD: SSDForce::getNonlinearity (registration.h:1646-1666) assembled as in
// FENonlinVectorOpInterface::applyAdd (FEOpInterface.h:1803-1853); every node sums the
// contributions of its (up to) four elements in the reference's element order.
// ================================================================================================
constexpr int GR_TX = 32, GR_TY = 8;  // nodes per workgroup tile
constexpr int GR_EX = GR_TX + 1, GR_EY = GR_TY + 1;

template <typename S, int DT>
__global__ void __launch_bounds__(256)
k_grad(const S *__restrict__ R, const S *__restrict__ T, const S *__restrict__ phi, S *__restrict__ DE, int n,
       double h, int width, const PairState *st, double lambda_default, int ntx, int nty, int npairs, int npsi) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * npairs);
  const int pair = wg / tiles, tile = wg - pair * tiles;
  if (st && !participates(st[pair])) return;
  const double lambda = st ? st[pair].lambda : lambda_default;

  __shared__ double cf[GR_EY * GR_EX][4][2];  // force contribution per element, dof, component
  __shared__ double ck[GR_EY * GR_EX][4][2];  // stiffness contribution

  const size_t nn = (size_t)n * n;
  const S *Rp = R + pair * nn, *Tp = T + pair * nn;
  const S *px = phi + pair * 2 * nn, *py = px + nn;
  const int x0 = (tile % ntx) * GR_TX, y0 = (tile / ntx) * GR_TY;
  const double vol = h * h;

  for (int e = threadIdx.x; e < GR_EX * GR_EY; e += 256) {
    const int ex = x0 - 1 + (e % GR_EX), ey = y0 - 1 + (e / GR_EX);
    if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
    double f0[4], f1[4], rn[4];
    for (int b = 0; b < 4; ++b) {
      const int g = l2g(ex, ey, b, n);
      f0[b] = ld(px, g);
      f1[b] = ld(py, g);
      rn[b] = ld(Rp, g);
    }
    double nl[4][2];
    if (DT == DT_CONS) {
      // reduceDeformations: sum over the n deformations psi_i (passed as T, 2 fields each) of
      // VariationOfConsistencyEnergyOp::getNonlinearity (AmbrosioTortorelli.h:755-799):
      // (Dpsi_i + I)^T (psi_i + offset) at the clipped deformed quad point, zero per clipped axis
      for (int q = 0; q < 4; ++q) {
        double ox = 0., oy = 0.;
        for (int b = 0; b < 4; ++b) {
          ox += f0[b] * bq(b, q);
          oy += f1[b] * bq(b, q);
        }
        const Hit hp = deform_qp(ex, ey, q, ox, oy, n, h, width);
        double a0 = 0., a1 = 0.;
        for (int k = 0; k < npsi; ++k) {
          const S *pk = T + (size_t)k * 2 * nn;
          double u4[4], w4[4];
          double vx = eval_img(pk, hp, n, u4), vy = eval_img(pk + nn, hp, n, w4);
          double j00 = 0., j01 = 0., j10 = 0., j11 = 0.;  // j[c][d] = d psi_c / d x_d
          for (int b = 0; b < 4; ++b) {
            const double gx = q1dx(b, hp.ly) / h, gy = q1dy(b, hp.lx) / h;
            j00 += gx * u4[b];
            j01 += gy * u4[b];
            j10 += gx * w4[b];
            j11 += gy * w4[b];
          }
          vx += ox;
          vy += oy;
          j00 += 1.;
          j11 += 1.;
          a0 += j00 * vx + j10 * vy;
          a1 += j01 * vx + j11 * vy;
        }
        nl[q][0] = hp.inx ? a0 : 0;
        nl[q][1] = hp.iny ? a1 : 0;
      }
    } else if (DT == DT_NCC) {
      // NormalizedCrossCorrelationEnergy::applyDerivative (registration.h:1974-1984): nodal weight
      // w = (Rn + E Tn) / v_t with Rn, Tn the normalised reference and warped template, then
      // CrossCorrelationForce (L1826-1862): -w(x_q) grad T at the unclipped deformed quad point,
      // zero where it leaves the grid
      const PairState &ps = st[pair];
      const bool divr = !(fabs(ps.ncc_rv - 0.0) < 1e-15), divt = !(fabs(ps.ncc_v - 0.0) < 1e-15);
      double w[4];
      for (int b = 0; b < 4; ++b) {
        double rh = rn[b] + (-ps.ncc_rm);
        if (divr) rh /= ps.ncc_rv;
        double th = fast_warp(Tp, n, h, ex + (b & 1), ey + ((b >> 1) & 1), f0[b], f1[b]) + (-ps.ncc_m);
        if (divt) th /= ps.ncc_v;
        double wb = rh;
        wb += ps.ncc_d * th;
        if (divt) wb /= ps.ncc_v;
        w[b] = wb;
      }
      for (int q = 0; q < 4; ++q) {
        double ox = 0., oy = 0.;
        for (int b = 0; b < 4; ++b) {
          ox += f0[b] * bq(b, q);
          oy += f1[b] * bq(b, q);
        }
        Hit hp;
        if (!transform_noclip(ex, ey, q, ox, oy, n, h, hp)) {
          nl[q][0] = nl[q][1] = 0.;
          continue;
        }
        double t4[4];
        (void)eval_img(Tp, hp, n, t4);
        double g0 = 0., g1 = 0.;
        for (int b = 0; b < 4; ++b) {
          double v0 = q1dx(b, hp.ly) / h, v1 = q1dy(b, hp.lx) / h;
          v0 *= t4[b];
          v1 *= t4[b];
          g0 += v0;
          g1 += v1;
        }
        double wq = 0.;
        for (int b = 0; b < 4; ++b) wq += w[b] * bq(b, q);
        const double sc = -wq;
        nl[q][0] = g0 * sc;
        nl[q][1] = g1 * sc;
      }
    } else {
      for (int q = 0; q < 4; ++q) {
        double ox = 0., oy = 0., r = 0.;
        for (int b = 0; b < 4; ++b) {
          ox += f0[b] * bq(b, q);
          oy += f1[b] * bq(b, q);
        }
        for (int b = 0; b < 4; ++b) r += rn[b] * bq(b, q);
        const Hit hp = deform_qp(ex, ey, q, ox, oy, n, h, width);
        double t4[4];
        const double td = eval_img(Tp, hp, n, t4);
        double g0 = 0., g1 = 0.;
        for (int b = 0; b < 4; ++b) {
          double v0 = q1dx(b, hp.ly) / h, v1 = q1dy(b, hp.lx) / h;
          v0 *= t4[b];
          v1 *= t4[b];
          g0 += v0;
          g1 += v1;
        }
        nl[q][0] = hp.inx ? g0 * (td - r) : 0;
        nl[q][1] = hp.iny ? g1 * (td - r) : 0;
      }
    }
    for (int dof = 0; dof < 4; ++dof) {
      double a0 = 0, a1 = 0;
      for (int q = 0; q < 4; ++q) {
        const double fct = bq(dof, q) * 0.25;
        a0 += fct * nl[q][0];
        a1 += fct * nl[q][1];
      }
      cf[e][dof][0] = vol * a0;
      cf[e][dof][1] = vol * a1;
    }
    // K_loc (1/6)[4 -1 -1 -2; -1 4 -2 -1; -1 -2 4 -1; -2 -1 -1 4] applied to the element's dofs
    const double kl[4][4] = {{4. / 6, -1. / 6, -1. / 6, -2. / 6},
                             {-1. / 6, 4. / 6, -2. / 6, -1. / 6},
                             {-1. / 6, -2. / 6, 4. / 6, -1. / 6},
                             {-2. / 6, -1. / 6, -1. / 6, 4. / 6}};
    for (int i = 0; i < 4; ++i) {
      double s0 = 0, s1 = 0;
      for (int j = 0; j < 4; ++j) {
        s0 += kl[i][j] * f0[j];
        s1 += kl[i][j] * f1[j];
      }
      ck[e][i][0] = s0;
      ck[e][i][1] = s1;
    }
  }
  __syncthreads();
  const int x = x0 + (threadIdx.x % GR_TX), y = y0 + (threadIdx.x / GR_TX);
  if (x >= n || y >= n) return;
  double F0 = 0, F1 = 0, K0 = 0, K1 = 0;
  for (int ey = y - 1; ey <= y; ++ey)
    for (int ex = x - 1; ex <= x; ++ex) {
      if (ex < 0 || ey < 0 || ex >= n - 1 || ey >= n - 1) continue;
      const int e = (ex - (x0 - 1)) + (ey - (y0 - 1)) * GR_EX;
      const int dof = (x - ex) + 2 * (y - ey);
      F0 += cf[e][dof][0];
      F1 += cf[e][dof][1];
      K0 += ck[e][dof][0];
      K1 += ck[e][dof][1];
    }
  S *o = DE + pair * 2 * nn;
  const int i = x + y * n;
  o[i] = static_cast<S>(F0 + lambda * K0);
  o[nn + i] = static_cast<S>(F1 + lambda * K1);
}

// ================================================================================================
// H1 metric: one V-cycle of the FD heat operator (LinearSmoothOp::apply, linearSmoothOp.cpp:591-612,
// MDMultigrid::solve L451-494, HeatMultigrid::jacobi L556-580) per field, 2 fields per pair.
// Down-leg kernel: x1 = rhs/diag (first Jacobi sweep from zero), x2 = x1 + (rhs - A x1)/diag,
// residual r = rhs - A x2, coarse rhs = 2x2 mean of r. Up-leg kernel: x = x2 + injected coarse
// correction, then two Jacobi sweeps. Tiles with a 2-node halo live in LDS.
// ================================================================================================
constexpr int MG_TX = 32, MG_TY = 16;
constexpr int MG_HX = MG_TX + 4, MG_HY = MG_TY + 4;
// FP32 fields: tiles of 64 x 16 nodes (4 per thread) on levels of 1024^2 and above -- 512 workgroups
// per field at 1024^2, resident at once, and a halo of 1.33x the tile instead of 1.41x -- and of
// 32 x 16 below, where a workgroup's latency, not the level's traffic, sets the pass's time.
constexpr int MG32_TY = 16, MG32_HY = MG32_TY + 4;
constexpr int MG32_TX_WIDE = 64, MG32_WIDE_FROM = 1024;

__device__ __forceinline__ void level_consts(const PairState *st, int pair, double sigma_default, int n, double &hs,
                                             double &diag) {
  const double sigma = st ? st[pair].sigma : sigma_default;
  const double tau = 0.5 * (sigma * sigma);  // LinearSmoothOp::setSigma (linearSmoothOp.h:50-52)
  const double s = (double)((n - 1) * (n - 1));
  hs = tau * s;
  diag = 1 + 4 * tau * s;
}

template <typename S>
__global__ void __launch_bounds__(256)
k_mg_down(const S *__restrict__ rhs, int negate, S *__restrict__ x2out, S *__restrict__ rhsc, int n,
          const PairState *st, double sigma_default, int ntx, int nty, int nfields) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * nfields);
  const int field = wg / tiles, tile = wg - field * tiles;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  double hs, diag;
  level_consts(st, pair, sigma_default, n, hs, diag);

  __shared__ double sr[MG_HY][MG_HX];   // rhs on tile+2
  __shared__ double s1[MG_HY][MG_HX];   // x1 on tile+2
  __shared__ double s2[MG_HY][MG_HX];   // x2 on tile+1
  __shared__ double sres[MG_TY][MG_TX]; // residual on tile
  const size_t nn = (size_t)n * n;
  const S *rp = rhs + field * nn;
  const int x0 = (tile % ntx) * MG_TX, y0 = (tile / ntx) * MG_TY;
  const int bx = x0 - 2, by = y0 - 2;  // LDS origin

  for (int i = threadIdx.x; i < MG_HX * MG_HY; i += 256) {
    const int lx = i % MG_HX, ly = i / MG_HX;
    const int gx = bx + lx, gy = by + ly;
    if (gx >= 0 && gy >= 0 && gx < n && gy < n) {
      double v = ld(rp, gx + gy * n);
      if (negate) v = -v;
      sr[ly][lx] = v;
      s1[ly][lx] = v / diag;  // 0 + (1.0 * (rhs - A 0)) / diag
    }
  }
  __syncthreads();
  auto get1 = [&](int gx, int gy) { return s1[gy - by][gx - bx]; };
  for (int i = threadIdx.x; i < MG_HX * MG_HY; i += 256) {
    const int lx = i % MG_HX, ly = i / MG_HX;
    const int gx = bx + lx, gy = by + ly;
    if (lx >= 1 && ly >= 1 && lx < MG_HX - 1 && ly < MG_HY - 1 && gx >= 0 && gy >= 0 && gx < n && gy < n) {
      const double ax = heat_apply(get1, gx, gy, n, hs);
      const double v = sr[ly][lx] - ax;
      s2[ly][lx] = s1[ly][lx] + 1.0 * v / diag;
    }
  }
  __syncthreads();
  auto get2 = [&](int gx, int gy) { return s2[gy - by][gx - bx]; };
  S *xo = x2out + field * nn;
  for (int i = threadIdx.x; i < MG_TX * MG_TY; i += 256) {
    const int tx = i % MG_TX, ty = i / MG_TX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx < n && gy < n) {
      const double ax = heat_apply(get2, gx, gy, n, hs);
      sres[ty][tx] = sr[ty + 2][tx + 2] - ax;
      xo[gx + gy * n] = static_cast<S>(s2[ty + 2][tx + 2]);
    }
  }
  __syncthreads();
  const int nc = n / 2;
  S *rc = rhsc + field * (size_t)nc * nc;
  for (int i = threadIdx.x; i < (MG_TX / 2) * (MG_TY / 2); i += 256) {
    const int cx = i % (MG_TX / 2), cy = i / (MG_TX / 2);
    const int gcx = x0 / 2 + cx, gcy = y0 / 2 + cy;
    if (gcx < nc && gcy < nc) {
      double v = 0;
      v += sres[2 * cy][2 * cx];
      v += sres[2 * cy][2 * cx + 1];
      v += sres[2 * cy + 1][2 * cx];
      v += sres[2 * cy + 1][2 * cx + 1];
      rc[gcx + gcy * nc] = static_cast<S>((1.0 / 4.0) * v);
    }
  }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_mg_up(const S *__restrict__ x2in, const S *__restrict__ xc, const S *__restrict__ rhs, int negate,
        S *__restrict__ xout, int n, const PairState *st, double sigma_default, int ntx, int nty,
        int nfields, const S *__restrict__ dold, double *__restrict__ nrm_part) {
  const int tiles = ntx * nty;
  const int wg = xcd_remap(blockIdx.x, tiles * nfields);
  const int field = wg / tiles, tile = wg - field * tiles;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  double hs, diag;
  level_consts(st, pair, sigma_default, n, hs, diag);

  __shared__ double sx[MG_HY][MG_HX];  // x on tile+2
  __shared__ double sr[MG_HY][MG_HX];  // rhs on tile+1
  __shared__ double s3[MG_HY][MG_HX];  // x3 on tile+1
  __shared__ double red[4 * NORM_TERMS];
  const size_t nn = (size_t)n * n;
  const int nc = n / 2;
  const S *x2p = x2in + field * nn, *rp = rhs + field * nn, *xcp = xc + field * (size_t)nc * nc;
  const int x0 = (tile % ntx) * MG_TX, y0 = (tile / ntx) * MG_TY;
  const int bx = x0 - 2, by = y0 - 2;

  for (int i = threadIdx.x; i < MG_HX * MG_HY; i += 256) {
    const int lx = i % MG_HX, ly = i / MG_HX;
    const int gx = bx + lx, gy = by + ly;
    if (gx >= 0 && gy >= 0 && gx < n && gy < n) {
      sx[ly][lx] = ld(x2p, gx + gy * n) + ld(xcp, (gx >> 1) + (gy >> 1) * nc);
      double v = ld(rp, gx + gy * n);
      sr[ly][lx] = negate ? -v : v;
    }
  }
  __syncthreads();
  auto getx = [&](int gx, int gy) { return sx[gy - by][gx - bx]; };
  for (int i = threadIdx.x; i < MG_HX * MG_HY; i += 256) {
    const int lx = i % MG_HX, ly = i / MG_HX;
    const int gx = bx + lx, gy = by + ly;
    if (lx >= 1 && ly >= 1 && lx < MG_HX - 1 && ly < MG_HY - 1 && gx >= 0 && gy >= 0 && gx < n && gy < n) {
      const double ax = heat_apply(getx, gx, gy, n, hs);
      const double v = sr[ly][lx] - ax;
      s3[ly][lx] = sx[ly][lx] + 1.0 * v / diag;
    }
  }
  __syncthreads();
  auto get3 = [&](int gx, int gy) { return s3[gy - by][gx - bx]; };
  S *xo = xout + field * nn;
  double acc[NORM_TERMS] = {0., 0., 0., 0.};  // <z,z>, <DE,z>, <DE,d_old>, <DE,DE> (DE = the raw input)
  for (int i = threadIdx.x; i < MG_TX * MG_TY; i += 256) {
    const int tx = i % MG_TX, ty = i / MG_TX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx < n && gy < n) {
      const double ax = heat_apply(get3, gx, gy, n, hs);
      const double v = sr[ty + 2][tx + 2] - ax;
      const S o = static_cast<S>(s3[ty + 2][tx + 2] + 1.0 * v / diag);
      xo[gx + gy * n] = o;
      if (nrm_part) {
        const double od = static_cast<double>(o), de = ld(rp, gx + gy * n);
        acc[0] += od * od;
        acc[1] += de * od;
        acc[2] += de * ld(dold + field * nn, gx + gy * n);
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

// Whole V-cycle below (and including) a level of at most 32x32 nodes in one workgroup, in LDS.
// With n equal to the registration level this is the complete metric; otherwise it is the coarse
// part of the cycle started by k_mg_down.
constexpr int MGC_MAXL = 5;
constexpr int MGC_OFF[8] = {0, 1, 5, 21, 85, 341, 1365, 5461};  // prefix sums of 4^l

template <typename S>
__global__ void __launch_bounds__(256)
k_mg_coarse(const S *__restrict__ rhs, int negate, S *__restrict__ xout, int L, const PairState *st,
            double sigma_default, int nfields, const S *__restrict__ dold, double *__restrict__ nrm_part) {
  const int field = blockIdx.x;
  const int pair = field >> 1;
  if (st && !participates(st[pair])) return;
  const double sigma = st ? st[pair].sigma : sigma_default;
  const double tau = 0.5 * (sigma * sigma);
  __shared__ double X[MGC_OFF[MGC_MAXL + 1]], B[MGC_OFF[MGC_MAXL + 1]], D[1024];
  __shared__ double red[4 * NORM_TERMS];
  const int n = 1 << L;
  const size_t nn = (size_t)n * n;
  const S *rp = rhs + field * nn;
  for (int i = threadIdx.x; i < (int)nn; i += 256) {
    double v = ld(rp, i);
    B[MGC_OFF[L] + i] = negate ? -v : v;
  }
  __syncthreads();

  // Jacobi sweep on level l: D = A x; x += (rhs - D)/diag
  auto sweep = [&](int l) {
    const int m = 1 << l;
    const double s = (double)((m - 1) * (m - 1));
    const double hs = tau * s, diag = 1 + 4 * tau * s;
    double *x = X + MGC_OFF[l];
    const double *b = B + MGC_OFF[l];
    auto get = [&](int gx, int gy) { return x[gx + gy * m]; };
    for (int i = threadIdx.x; i < m * m; i += 256) D[i] = heat_apply(get, i & (m - 1), i >> l, m, hs);
    __syncthreads();
    for (int i = threadIdx.x; i < m * m; i += 256) {
      const double v = b[i] - D[i];
      x[i] += 1.0 * v / diag;
    }
    __syncthreads();
  };

  for (int l = L; l >= 1; --l) {
    const int m = 1 << l;
    double *x = X + MGC_OFF[l];
    for (int i = threadIdx.x; i < m * m; i += 256) x[i] = 0;
    __syncthreads();
    sweep(l);
    sweep(l);
    const double s = (double)((m - 1) * (m - 1));
    const double hs = tau * s;
    auto get = [&](int gx, int gy) { return x[gx + gy * m]; };
    for (int i = threadIdx.x; i < m * m; i += 256) D[i] = B[MGC_OFF[l] + i] - heat_apply(get, i & (m - 1), i >> l, m, hs);
    __syncthreads();
    const int mc = m / 2;
    for (int i = threadIdx.x; i < mc * mc; i += 256) {
      const int cx = i & (mc - 1), cy = i >> (l - 1);
      double v = 0;
      v += D[2 * cx + 2 * cy * m];
      v += D[2 * cx + 1 + 2 * cy * m];
      v += D[2 * cx + (2 * cy + 1) * m];
      v += D[2 * cx + 1 + (2 * cy + 1) * m];
      B[MGC_OFF[l - 1] + i] = (1.0 / 4.0) * v;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {  // 1x1 level: 100 sweeps of x += (rhs - x) / 1
    double x = 0;
    const double b = B[0];
    for (int it = 0; it < 100; ++it) {
      const double v = b - x;
      x += 1.0 * v / 1.0;
    }
    X[0] = x;
  }
  __syncthreads();
  for (int l = 1; l <= L; ++l) {
    const int m = 1 << l, mc = m / 2;
    double *x = X + MGC_OFF[l];
    const double *xcs = X + MGC_OFF[l - 1];
    for (int i = threadIdx.x; i < m * m; i += 256) x[i] += xcs[((i & (m - 1)) >> 1) + ((i >> l) >> 1) * mc];
    __syncthreads();
    sweep(l);
    sweep(l);
  }
  S *xo = xout + field * nn;
  double acc[NORM_TERMS] = {0., 0., 0., 0.};  // <z,z>, <DE,z>, <DE,d_old>, <DE,DE> (DE = the raw input)
  for (int i = threadIdx.x; i < (int)nn; i += 256) {
    const S o = static_cast<S>(X[MGC_OFF[L] + i]);
    xo[i] = o;
    if (nrm_part) {
      const double od = static_cast<double>(o), de = ld(rp, i);
      acc[0] += od * od;
      acc[1] += de * od;
      acc[2] += de * ld(dold + field * nn, i);
      acc[3] += de * de;
    }
  }
  if (nrm_part) {
    block_sum_n<NORM_TERMS>(acc, red);
    if (threadIdx.x == 0)
      for (int t = 0; t < NORM_TERMS; ++t) nrm_part[(size_t)field * NORM_TERMS + t] = acc[t];
  }
}

// ================================================================================================
// Update phi += tau * dir for pairs whose search accepted a step (GradientDescentBase::applySingle
// L382-385). Along the CG direction the kernel forms d = z + beta d_old (search_dir) and stores it
// in place of d_old: it is the next iteration's old direction.
// ================================================================================================
constexpr int UPD_PER_THREAD = 4;

template <typename S>
__global__ void __launch_bounds__(256)
k_update(S *__restrict__ phi, S *__restrict__ d, const S *__restrict__ z, int n, const PairState *st, int nblk_pair,
         int npairs, int slot) {
  const int total = nblk_pair * npairs;
  const int wg = xcd_remap(blockIdx.x, total);
  const int pair = wg / nblk_pair, blk = wg - pair * nblk_pair;
  const double tau = st[pair].step_tau[slot];
  if (tau == 0) return;  // no step accepted in this iteration
  const bool cg = st[pair].step_dir[slot] == DIR_CG, use_old = st[pair].cg_old != 0;
  const double beta = st[pair].beta;
  const size_t off = pair * 2 * (size_t)n * n;
  S *p = phi + off, *dp = d + off;
  const S *zp = z + off;
  const int m = 2 * n * n;
  const int base = blk * 256 * UPD_PER_THREAD;
  // every load first (indices clamped into the field), then the stores: the loads of all
  // UPD_PER_THREAD nodes are in flight together instead of one node's at a time
  S zv[UPD_PER_THREAD], ov[UPD_PER_THREAD], pv[UPD_PER_THREAD];
#pragma unroll
  for (int k = 0; k < UPD_PER_THREAD; ++k) {
    const int i = min(base + k * 256 + (int)threadIdx.x, m - 1);
    zv[k] = zp[i];
    ov[k] = dp[i];
    pv[k] = p[i];
  }
#pragma unroll
  for (int k = 0; k < UPD_PER_THREAD; ++k) {
    const int i = base + k * 256 + threadIdx.x;
    if (i < m) {
      const S dv = search_dir_val<S>(zv[k], ov[k], cg, use_old, beta);
      if (cg) dp[i] = dv;
      p[i] = step_at<S>(pv[k], tau, dv);
    }
  }
}

// ================================================================================================
// Grid transfer and warps
// ================================================================================================

// CellCenteredRestrictOp::apply (cellCenteredGrid.h:165-180): 2x2 mean, one coarse node per thread.
template <typename S>
__global__ void k_restrict(const S *__restrict__ fine, S *__restrict__ coarse, int n, int nfields) {
  const int nc = n / 2;
  const size_t cnt = (size_t)nc * nc;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt * nfields) return;
  const int f = (int)(i / cnt);
  const int j = (int)(i - (size_t)f * cnt);
  const int cx = j % nc, cy = j / nc;
  const S *fp = fine + (size_t)f * n * n;
  double v = 0;
  v += ld(fp, 2 * cx + 2 * cy * n);
  v += ld(fp, 2 * cx + 1 + 2 * cy * n);
  v += ld(fp, 2 * cx + (2 * cy + 1) * n);
  v += ld(fp, 2 * cx + 1 + (2 * cy + 1) * n);
  coarse[i] = static_cast<S>((1.0 / 4.0) * v);
}

// CellCenteredProlongOp::apply (cellCenteredGrid.h:188-210): injection n -> 2n.
template <typename S>
__global__ void k_prolong(const S *__restrict__ coarse, S *__restrict__ fine, int nc, int nfields) {
  const int n = 2 * nc;
  const size_t cnt = (size_t)n * n;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt * nfields) return;
  const int f = (int)(i / cnt);
  const int j = (int)(i - (size_t)f * cnt);
  const int x = j % n, y = j / n;
  fine[i] = coarse[(size_t)f * nc * nc + (x >> 1) + (y >> 1) * nc];
}

// ScalarArray<QC_2D>::resampleFrom (scalarArray.h:308-331, 1052-1060); coord[k] = k * (nc-1)/(n-1)
// evaluated in long double on the host and rounded to double.
template <typename S>
__global__ void k_resample(const S *__restrict__ src, S *__restrict__ dst, int nc, int n,
                           const double *__restrict__ coord, int nfields) {
  const size_t cnt = (size_t)n * n;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt * nfields) return;
  const int f = (int)(i / cnt);
  const int j = (int)(i - (size_t)f * cnt);
  dst[i] = static_cast<S>(interp(src + (size_t)f * nc * nc, nc, coord[j % n], coord[j / n]));
}

// ScalarArray<QC_2D>::resampleFrom of a w x h source onto n x n nodes: node (x, y) takes the source
// interpolated at (cx[x], cy[y]).
template <typename S>
__global__ void k_resample_rect(const S *__restrict__ src, int w, int h, S *__restrict__ dst, int n,
                                const double *__restrict__ cx, const double *__restrict__ cy) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * n) return;
  const int x = (int)(i % n), y = (int)(i / n);
  dst[i] = static_cast<S>(interp_wh(src, w, h, cx[x], cy[y]));
}

// DeformImage (deformations.h:62-101; ExtendWithConstant, optional nearest-neighbour rounding),
// FastDeformImage (L140-162; clamp) and the NaN-extended variant of L226-256 (mode 2).
// img: nimg images; phi: one field per image (or one shared field if phi_shared).
template <typename S>
__global__ void k_warp(const S *__restrict__ img, const S *__restrict__ phi, S *__restrict__ out, int n, int nimg,
                       int phi_shared, int mode, double ext, int nearest) {
  const size_t cnt = (size_t)n * n;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt * nimg) return;
  const int f = (int)(i / cnt);
  const int j = (int)(i - (size_t)f * cnt);
  const int x = j % n, y = j / n;
  const double h = 1.0 / (double)(n - 1);
  const S *ph = phi + (phi_shared ? 0 : (size_t)f * 2 * cnt);
  double dx = x + ld(ph, j) / h, dy = y + ld(ph, cnt + j) / h;
  const S *im = img + (size_t)f * cnt;
  double v;
  if (mode == 1) {  // clamp
    const double hi = (double)(n - 1);
    dx = dx < 0 ? 0 : (dx > hi ? hi : dx);
    dy = dy < 0 ? 0 : (dy > hi ? hi : dy);
    v = interp(im, n, dx, dy);
  } else {
    bool inside = true;
    const double hi = (double)(n - 1);
    if (dx < 0.0 || dx > hi || (mode == 0 && isnan(dx))) inside = false;
    if (dy < 0.0 || dy > hi || (mode == 0 && isnan(dy))) inside = false;
    if (nearest && mode == 0) {
      dx = floor(dx + 0.5);
      dy = floor(dy + 0.5);
    }
    v = inside ? interp(im, n, dx, dy) : (mode == 2 ? __builtin_nan("") : ext);
  }
  out[i] = static_cast<S>(v);
}

}  // namespace msg
