/*
 * ms_oracle.h — FP64 CPU restatement of match-series' registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so, and only as the checker (or the timed CPU baseline). The product never calls it.
 *
 * Conventions (shared with the reference, QuocMesh):
 *   - a level-l grid has N = 2^l nodes per side, h = 1/(N-1) (rectangularGrid.h:50-57);
 *   - node (x, y) is stored at x + y*N (configurators.h:889-900);
 *   - a displacement field phi is two consecutive N*N arrays (x component, then y component), in
 *     domain units: the pixel offset is phi/h (deformations.h:85).
 * Every function cites the reference code it restates.
 */
#ifndef MS_ORACLE_H
#define MS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* --- grid transfer (cellCenteredGrid.h:157-210, scalarArray.h:308-331,1052-1060) --- */
void msor_restrict(const double *fine, int N, double *coarse);        /* N -> N/2, 2x2 mean      */
void msor_prolong(const double *coarse, int n, double *fine);         /* n -> 2n, injection      */
void msor_resample(const double *src, int n, double *dst, int N);     /* n -> N, bilinear        */

/* --- images (scalarArray.h:499-529, deformations.h:62-101,140-162) --- */
double msor_interpolate(const double *img, int N, double X, double Y);
void msor_deform_image(const double *img, const double *phi, int N, double ext, int nearest, double *out);
void msor_fast_deform_image(const double *img, const double *phi, int N, double *out);

/* --- energy terms (registration.h:1607-1615,1646-1666; FEOpInterface.h:1803-1853,2586-2613) --- */
double msor_ssd_energy(const double *R, const double *T, const double *phi, int N);
void msor_ssd_force(const double *R, const double *T, const double *phi, int N, double *out);
void msor_stiff_apply(const double *u, int N, double *out);           /* one component, K u      */
double msor_reg_energy(const double *phi, int N);                     /* 1/2 sum_c phi_c K phi_c */
void msor_mass_apply(const double *u, int N, double *out);            /* one component, M u      */
/* NCC data term (deformationModel 0; registration.h:1909-2018) */
double msor_ncc_normalize(const double *img, int N, double *out);     /* returns sqrt(u.Mu)      */
double msor_ncc_energy(const double *R, const double *T, const double *phi, int N);
void msor_ncc_force(const double *R, const double *T, const double *phi, int N, double *out);

/* --- H1 metric: one FD heat-equation multigrid V-cycle (linearSmoothOp.cpp:11-147,451-612) --- */
void msor_heat_apply(const double *u, int N, double tau, double *out);
void msor_vcycle(const double *d, int N, double sigma, double *out);

/* --- one level of the gradient flow (gradientDescent.h:200-236,295-464; ArmijoSearch.h:118-200) ---
 * phi (2*N*N) is updated in place. Traces (may be NULL) are written per Armijo call
 * (tau, sigma at the call, energy evaluations of the call) and per iteration (energy after the
 * step, energy_trace[0] = initial energy); trace_cap bounds every trace. Returns iterations. */
/* data term: deformationModel 22 / deformationModel 0 / the seismic line-wise x-shift NCC */
enum { MSOR_SSD = 0, MSOR_NCC = 1, MSOR_XSHIFT = 2 };

typedef struct {
  double lambda;      /* regularisation weight of this level            */
  int max_iter;       /* maxGDIterations                                */
  double stop_eps;    /* stopEpsilon                                    */
  int data_term;      /* MSOR_SSD, MSOR_NCC or MSOR_XSHIFT              */
  double y_reg;       /* MSOR_XSHIFT: yRegFactor                        */
} msor_level_params;

/* --- seismic x-shift (FELineShiftRegisEnergy, registration.h:2137-2193; findXShift,
 * matchSeries.h:805-868): phix is the N*N x component; the y component of the solver's field stays
 * zero. Energies unweighted (E = e_data + lambda e_reg); the gradient includes lambda A phix. --- */
void msor_xshift_energy(const double *R, const double *T, const double *phix, int N, double yreg, double *e_data,
                        double *e_reg);
void msor_xshift_gradient(const double *R, const double *T, const double *phix, int N, double yreg, double lambda,
                          double *out);

int msor_level_solve(const double *R, const double *T, double *phi, int N, const msor_level_params *p,
                     double *energy_trace, double *tau_trace, double *sigma_trace, int *evals_trace,
                     int trace_cap, int *n_armijo_calls, double *final_energy);

/* --- reduceDeformations (matchSeries.h:678-704; AmbrosioTortorelli.h:680-800) ---
 * psi: n displacement fields (2*N*N each); phi (out): the reduced deformation. Returns iterations. */
double msor_consistency_energy(const double *psi, const double *phi, int N);
void msor_consistency_force(const double *psi, const double *phi, int N, double *out);
int msor_reduce_deformations(const double *psi, int n, int N, int max_iter, double stop_eps, double *phi,
                             double *final_energy);

/* --- input normalisation and image pyramid (registration.h:331-367; vec.h:864-870,1084-1089) ---
 * pct < 0 means "enhanceContrastSaturationPercentage" absent. */
void msor_normalize(double *img, int count, double pct, int min_to_zero);

/* --- deformation composition with harmonic extension (deformations.h:364-404,499-520) ---
 * result = phi1(x + phi2(x)) + phi2(x) where defined, extended by a Dirichlet-Laplace PCG solve.
 * result holds the PCG start value on entry (the reference passes its output array, which holds
 * the previous accumulated deformation). */
void msor_compose(const double *phi1, const double *phi2, int N, double *result);

/* --- reconstruction statistics (matchSeries.h:406-444; multiVector.cpp:99-117; vec.cpp:933-949) ---
 * stack: F arrays of NN values, +Inf marks a missing sample. */
void msor_series_stats(const double *stack, int F, int NN, double fill, double *mean, double *count,
                       double *median);

/* --- one stage of series matching (matchSeries.h:87-269,353-451,1249-1320) ---
 * templates: n raw images of N*N values; reference: raw N*N image (stage 1: the first frame,
 * later stages: the previous stage's median). Images are normalised for registration (unless
 * dont_normalize) and averaged raw, as the reference does.
 * pairwise_in: NULL, or n fields of 2*N*N whose entry i (i >= 1) replaces the pairwise solve of
 *   template i ("reuseStage1Results": the previous stage's pairwise deformation of template i-1).
 * deformations (out): n fields, the deformation each template is averaged with.
 * pairwise_out (out, may be NULL): n fields, each template's pairwise deformation.
 * energies (out): n final energies; mean/count/median (out): N*N each. */
typedef struct {
  int start_level, stop_level, refine_start_level, refine_stop_level, precision_level;
  double lambda, lambda_factor;       /* lambda_factor <= 0: "lambdaFactor" absent            */
  int max_iter;
  double stop_eps;
  double saturation_pct;              /* < 0: absent                                          */
  int normalize_min_to_zero;
  int dont_normalize;                 /* dontNormalizeInputImages                             */
  int resample_prolong;               /* resampleInsteadOfProlongateDeformation               */
  int dont_accumulate;                /* dontAccumulateDeformation                            */
  int data_term;                      /* MSOR_SSD or MSOR_NCC                                 */
  int reduce_deformations;            /* reduceDeformations                                   */
  int reduce_max_iter;
  double reduce_stop_eps;
  const double *pair0_init;           /* NULL or template 0's starting deformation            */
  int corr_init;                      /* useCorrelationToInitTranslation                      */
  int max_corr_shift;                 /* maxCorrShift                                         */
  int calc_inverse;                   /* calcInverseDeformation                               */
  double *inverse_out;                /* NULL or n fields: the inverse solves' deformations   */
  int alt_start_level;                /* altStartLevel (0: none, or equal to start_level)     */
  int energy_as_alt_criterion;        /* energyAsAltCriterion                                 */
  double *first_try_out;              /* NULL or n fields: the start_level solves' results    */
} msor_series_params;

/* Convolution::phaseCorrelation (quoc/convolution.h:1049-1114): the inverse FFT of
 * FFT(a) * conj(FFT(b)), scaled to unit modulus per frequency when `normalise`, times 1/N^2. */
void msor_phase_correlation(const double *a, const double *b, int N, int normalise, double *out);
/* PhaseCorrelationRegistration::registerImages (quoc/paramReg.h:2057-2096): signed integer shift,
 * x then y. */
void msor_register_translation(const double *reference, const double *templ, int N, int max_shift, int *shift);

/* ScalarArray<QC_2D>::resampleFrom of a w x h source onto N x N (scalarArray.h:308-331) */
void msor_resample_rect(const double *src, int w, int h, double *dst, int N);

/* qc::approxInverseDeformation (quoc/deformations.h:341-355): the identity pushed through phi by
 * TransformFunction<QC_2D>::transform (quoc/auxiliary.cpp:33-176: two triangles per element, x-fastest
 * element order, the first triangle covering a node wins), minus the identity; nodes no triangle
 * covers get the midpoint of the covered values' range. phi, inv: (2, N, N). */
void msor_approx_inverse(const double *phi, int N, double *inv);

int msor_match_series(const double *templates, int n, const double *reference, int N,
                      const msor_series_params *p, const double *pairwise_in,
                      double *deformations, double *pairwise_out, double *energies,
                      double *mean, double *count, double *median, double *reduced);

#ifdef __cplusplus
}
#endif
#endif
