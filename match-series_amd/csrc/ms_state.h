// Per-pair scalar state of a batched level solve. Lives in device memory; only the control kernel
// writes it (one workgroup per pair), every other kernel reads it. All scalars are FP64.
#pragma once

namespace msg {

enum ArmijoPhase { AP_IDLE = 0, AP_FIRST = 1, AP_GROW = 2, AP_SHRINK = 3 };
enum DirMode { DIR_CG = 0, DIR_GRAD = 1, DIR_NONE = 2 };
enum { MAX_CAND = 4 };
// deformationModel 22 / deformationModel 0 / reduceDeformations' consistency energy / the seismic
// row-wise x-shift NCC (matchSeismicSeries)
enum DataTerm { DT_SSD = 0, DT_NCC = 1, DT_CONS = 2, DT_XSHIFT = 3 };

struct PairState {
  // gradient-flow state (GradientDescentBase::applySingle)
  double energy;      // E at the current position ("energy", also the Armijo f)
  double energy_new;  // E after the last step
  double ereg_new;    // E_reg (unweighted) after the last step
  double tau;         // step size carried to the next Armijo search
  double sigma;       // filter width of the H1 metric
  double beta;        // Fletcher-Reeves coefficient of this iteration
  double zz_old;      // |old smoothed gradient|^2
  double Df, Dfz;     // <DE, d>, <DE, z>
  double lambda;      // regularisation weight of the level
  double stop_eps;
  double energy_start, ereg_start;  // E and E_reg (unweighted) at the start of the level
  // Armijo machine (getTimestepWidthWithArmijoLineSearch)
  double a_tau;       // tau of the sequential algorithm
  double a_Df;        // slope of the active direction
  double cand[MAX_CAND];
  double cand_E[MAX_CAND], cand_R[MAX_CAND];
  double last_tau[2], last_E[2], last_R[2];  // the last two consumed evaluations
  // NCC data term: per evaluation the data energy and the mass-weighted mean / deviation of the
  // warped template; the reference's mean / deviation of the level; the values at the current
  // position (what the variation needs)
  double cand_D[MAX_CAND], cand_M[MAX_CAND], cand_V[MAX_CAND];
  double last_D[2], last_M[2], last_V[2];
  double ncc_rm, ncc_rv, ncc_rb;  // reference: M-mean, deviation, 1'M(r - mean) (0 up to rounding)
  double ncc_c;                   // shift of the template moments (a recent template mean)
  double ncc_d, ncc_m, ncc_v;
  double dd;          // |DE|^2 at the current position (gradient-norm stopping)
  int sigma_fixed;    // 1: filter width fixed (plain H1GradientDescent), 0: the automatic schedule
  int grad_stop;      // 1: stop on |DE| <= stopEpsilon, 0: on the energy decrease
  double step_tau[2];  // per step slot: tau of the accepted step (0: none)
  double tau_first;   // result of the CG-direction search this iteration (trace)
  int ncand;          // candidates to evaluate in the next energy pass (0: none)
  int a_phase;
  int dirmode;        // direction the energy pass / update uses
  int step_dir[2];
  int iter, max_iter;
  int active;         // still iterating at this level
  int searching;      // Armijo search in progress
  int olddir_zero;    // the previous search fell back: treat the old CG direction as zero
  int cg_old;         // this iteration's CG direction is z + beta d_old (0: z; the old direction is zero)
  int evals;          // evaluations consumed by this iteration's searches (reference count)
  int passes;         // energy passes launched for this iteration
  int total_evals, total_passes;
  // FP32 energy passes: a step decision whose margin lies within the pass's error bound is taken on
  // values evaluated again in FP64 arithmetic at the stored positions
  int exact;          // the planned pass evaluates in FP64 arithmetic (candidates: the step, and 0 for f)
  int cand_ex;        // the candidates' values come from such a pass
  int last_ex[2];     // the same for the last two consumed evaluations
  int f_ex;           // `energy` (the Armijo f) comes from such a pass
};

// Per-iteration trace record (optional).
struct IterTrace {
  double energy;    // energy after the step
  double ereg;      // unweighted regularisation energy after the step
  double tau;       // final step size
  double tau_first; // step size of the CG-direction search (0 if it fell back)
  double sigma;     // filter width used for the step
  double evals;     // reference-equivalent energy evaluations
  double passes;    // energy passes
  double beta;
};

}  // namespace msg
