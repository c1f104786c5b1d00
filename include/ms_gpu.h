/*
 * ms_gpu.h — C ABI of the MI355X (gfx950) implementation of match-series' registration hot path.
 *
 * Library: match-series_amd/lib/libmsgpu.so. Every function returns MSG_OK (0) or an error code;
 * msg_last_error() describes the last failure of the calling thread. No function throws.
 *
 * Conventions (those of the reference, QuocMesh):
 *   - a level-l grid has n = 2^l nodes per side, h = 1/(n-1); node (x, y) is stored at x + y*n;
 *   - an image is n*n values; a displacement field is 2*n*n values (x component, then y), in
 *     domain units (pixel offset = value / h);
 *   - a batch of k images/fields is stored back to back.
 * Host arrays are float64 (the reference's RealType). Device arrays carry a dtype: MSG_F64 stores and
 * computes in FP64 in the reference's operation order. MSG_F32 stores FP32; with the SSD data term the
 * warp, the stencils and the V-cycle compute in FP32 (energy sums, slopes, norms and every step
 * decision stay FP64), the NCC and consistency energies in FP64 on the FP32 values.
 *
 * Each entry point names the reference interface it replaces (file:line, QuocMesh sources under
 * quocmesh/modules/ and quocmesh/projects/electronMicroscopy/).
 */
#ifndef MS_GPU_H
#define MS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum { MSG_OK = 0, MSG_ERR_ARG = 1, MSG_ERR_HIP = 2, MSG_ERR_STATE = 3, MSG_ERR_ALLOC = 4 };
enum { MSG_F32 = 0, MSG_F64 = 1 };
/* Data terms: SSD (matchSeries deformationModel 22, SSDRegistrationConfigurator), normalised
 * cross-correlation (deformationModel 0, NCCRegistrationConfigurator; registration.h:1766-1781) and the
 * seismic row-wise x-shift (matchSeismicSeries: FELineShiftRegisEnergy, registration.h:2137-2193, with
 * findXShift's regulariser K_xx + yRegFactor K_yy, matchSeries.h:805-868; the y component of the
 * deformation stays zero). */
enum { MSG_SSD = 0, MSG_NCC = 1, MSG_XSHIFT = 3 };

/* Description of the last error raised on the calling thread ("" if none). */
const char *msg_last_error(void);
/* Library version string. */
const char *msg_version(void);

/* ------------------------------------------------------------------------------------------------
 * Registration context: a batch of `npairs` independent (reference, template, deformation) triples
 * with image pyramids up to level `max_level` — the state of
 * qc::RegistrationMultilevelDescentInterface (quoc/registration.h:422-672) and
 * qc::StandardRegistrationMultilevelDescent (registration.h:1072-1132), one per pair.
 * ---------------------------------------------------------------------------------------------- */
typedef struct msg_ctx msg_ctx;
/* Streams: a context works on its own non-blocking HIP stream. Device arrays handed to a context
 * call (or to msg_match_series) must be complete when the call is made — synchronise the stream
 * that produced them — and the library has finished with every argument when the call returns. */

/* RegistrationMultilevelDescentInterfaceBase(Parser) with precisionLevel = max_level
 * (registration.h:263-274); allocates every device buffer the solve needs. */
int msg_ctx_create(msg_ctx **ctx, int max_level, int npairs, int dtype);
int msg_ctx_destroy(msg_ctx *ctx);

/* setReference / setTemplate (registration.h:310-319): host float64 image at max_level; the pyramid
 * is rebuilt by 2x2 restriction (MultilevelArray::levRestrict, multilevelArray.h:99-110). */
int msg_set_reference(msg_ctx *ctx, int pair, const double *img);
int msg_set_template(msg_ctx *ctx, int pair, const double *img);
/* Same for all pairs at once from device memory (npairs images of dtype `dtype`). */
int msg_set_images_device(msg_ctx *ctx, const void *d_ref, const void *d_tpl, int dtype);

/* setTransformation / getTransformation / setTransformationToZero (registration.h:549-576):
 * host float64 field at max_level; setting restricts it to every level. */
int msg_set_deformation(msg_ctx *ctx, int pair, const double *phi);
int msg_get_deformation(msg_ctx *ctx, int pair, double *phi);
int msg_set_deformation_zero(msg_ctx *ctx, int pair);
/* The field of one level alone (2 x 2^level x 2^level float64, x then y), the other levels left as
 * they are: the deformation of MultiDimMultilevelArray::setCurLevel(level) that descentOnCurrentGrid
 * works on (registration.h:1092-1112). */
int msg_set_deformation_level(msg_ctx *ctx, int pair, int level, const double *phi);
int msg_get_deformation_level(msg_ctx *ctx, int pair, int level, double *phi);
/* Device copy of all pairs' max_level fields (dtype of the context), in or out. */
int msg_set_deformations_device(msg_ctx *ctx, const void *d_phi);
int msg_get_deformations_device(msg_ctx *ctx, void *d_phi);

/* Parameters of the gradient flow (StandardRegistration::setSettingsFromParser,
 * registration.h:1059-1066; getRegularizationWeight, L523-526). */
typedef struct {
  double lambda;          /* "lambda"                                                         */
  double lambda_factor;   /* "lambdaFactor"; <= 0: absent                                     */
  int lambda_ref_level;   /* "stopLevel" the lambda factor is anchored to                     */
  int max_iter;           /* "maxGDIterations" (reference default 1000)                      */
  double stop_eps;        /* "stopEpsilon" (reference default 0)                             */
  int resample_prolong;   /* "resampleInsteadOfProlongateDeformation"                         */
  int spec;               /* Armijo candidates per energy pass, 1..4 (0: default 3; a pass   */
                          /* never plans more than 3, so 4 acts as 3)                          */
} msg_solve_params;

/* Per-iteration record of a level solve (the content of energy_<LL>.txt plus step data). */
typedef struct {
  double energy, ereg, tau, tau_first, sigma, evals, passes, beta;
} msg_iter_trace;

/* One level of gradient flow for every pair (descentOnCurrentGrid, registration.h:1092-1112 ->
 * StandardRegistration::findTransformation L967-1000 -> GradientDescentBase::applySingle,
 * aol/gradientDescent.h:295-464). energies/iterations: npairs outputs (may be NULL).
 * trace: NULL or npairs * trace_cap records. */
int msg_solve_level(msg_ctx *ctx, int level, double lambda, int max_iter, double stop_eps, int spec,
                    double *energies, int *iterations, msg_iter_trace *trace, int trace_cap);

/* solveAndProlongToMaxDepth(start, stop) for every pair (registration.h:462-467 over
 * MultilevelDescentInterface::solve, quoc/quocDescent.h:83-105); energies = getEnergyOfLastSolution. */
int msg_solve(msg_ctx *ctx, int start_level, int stop_level, const msg_solve_params *p, double *energies);

/* How a level solve is driven (no counterpart in the reference; results are identical):
 * 0: the host reads the per-pair flags after every Armijo energy pass;
 * 1: device-driven — each control phase is its own one-workgroup-per-pair kernel after the pass it
 *    reduces, the host checks flags every 8 iterations; 2: as 1 with the 8 iterations replayed as
 *    one hipGraph. Default 0.
 * msg_set_default_mode applies to contexts created later and to msg_match_series. */
int msg_ctx_set_mode(msg_ctx *ctx, int mode);
/* The registration configurator of the context's solves (StandardRegistrationMultilevelDescent's
 * RegistrationConfiguratorType, matchSeries.cpp:17-21 and 194-197): MSG_SSD (default) or MSG_NCC; or
 * the x-shift registration of XShiftRegistrationMultilevelDescent (MSG_XSHIFT, rows of at most 2048
 * nodes; matchSeries.h:925-982). */
int msg_ctx_set_data_term(msg_ctx *ctx, int data_term);
/* "yRegFactor" of the x-shift regulariser (default 1). */
int msg_ctx_set_y_reg_factor(msg_ctx *ctx, double y_reg_factor);
int msg_set_default_mode(int mode);
/* FP32 fields with the SSD data term compute the energy passes in FP32 arithmetic. A step decision
 * of the Armijo search (G >= 1/2, f >= E; ArmijoSearch.h:148-190) whose margin lies within the
 * passes' error bound is then taken on both values evaluated again in FP64 arithmetic at the stored
 * positions (on: 1, the default), so that FP32 rounding alone does not flip it; 0 decides on the
 * FP32-arithmetic values. No counterpart in the reference; applies to contexts created later and to
 * msg_match_series. */
int msg_set_default_exact_decisions(int on);

/* Totals of the last msg_solve / msg_solve_level: iterations, reference-equivalent energy
 * evaluations and energy passes summed over pairs and levels. */
int msg_last_counts(msg_ctx *ctx, long long *iterations, long long *evals, long long *passes);

/* Device pointer of a pair's level-l image or field inside the context (for zero-copy callers). */
int msg_ctx_buffer(msg_ctx *ctx, int which /* 0 ref, 1 tpl, 2 phi */, int level, void **ptr);

/* ------------------------------------------------------------------------------------------------
 * Operator surface on device arrays (batched). stream: hipStream_t or NULL (default stream).
 * ---------------------------------------------------------------------------------------------- */

/* SSDEnergy (registration.h:1584-1616) and DisplacementLengthEnergy (quoc/hyperelastic.h:168-190):
 * per pair E_SSD(phi) and 1/2 sum_c phi_c^T K phi_c (host outputs, batch values each). */
int msg_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ssd,
               double *e_reg, void *stream);
/* msg_energy without any allocation (no counterpart in the reference). For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. Same kernels and summation order as msg_energy, so the results are bit-identical.
 * Its pair states and partial sums live in a fixed 32 MiB device array that each GPU holds until the
 * process ends; the call fails with MSG_ERR_ARG when batch x n^2 needs more (about
 * (8 batch - 6) ceil((n-1)/64) ceil((n-1)/r) doubles, r = 4 for FP64 fields, 16 for FP32 fields of
 * 1024^2 and above, 8 below: e.g. 32 pairs of 2048^2 in FP64, 128 in FP32). Every call on one
 * GPU shares that array, so calls that overlap on one GPU (several threads or streams) race;
 * serialise them. The call synchronises `stream` before returning. */
int msg_energy_noalloc(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ssd,
                       double *e_reg, void *stream);

/* SSDForce (registration.h:1622-1667) + lambda * K phi (DirichletRegularizationConfigurator,
 * registration.h:677-710): the variation DE of the registration energy. */
int msg_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch,
                 double lambda, void *stream);
/* NormalizedCrossCorrelationEnergy (registration.h:1909-1990; FastDeformImage warp, assembled mass
 * matrix) per pair: e_ncc (and e_reg as msg_energy); its variation VariationOfNormalizedCross-
 * CorrelationEnergy (L1998-2018, CrossCorrelationForce L1826-1862) plus lambda K phi. */
int msg_ncc_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double *e_ncc,
                   double *e_reg, void *stream);
int msg_ncc_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch,
                     double lambda, void *stream);
/* The x-shift energy of findXShift (projects/electronMicroscopy/matchSeries.h:805-868) per pair, phi
 * laid out as for msg_energy with only its x component read: e_ncc = FELineShiftRegisEnergy::apply
 * (registration.h:2146-2156, the 1-D NCC of every row, rows starting with NaN left out, divided by
 * the number of rows) and e_reg = 1/2 phi_x^T (K_xx + y_reg_factor K_yy) phi_x (matchSeries.h:821-828).
 * The variation: DE_x = FELineShiftRegisEnergy::applyDerivative (registration.h:2161-2175) + lambda
 * (K_xx + y_reg_factor K_yy) phi_x, DE_y = 0. Rows of at most 2048 nodes. */
int msg_xshift_energy(int dtype, const void *R, const void *T, const void *phi, int n, int batch, double y_reg_factor,
                      double *e_ncc, double *e_reg, void *stream);
int msg_xshift_gradient(int dtype, const void *R, const void *T, const void *phi, void *DE, int n, int batch,
                        double lambda, double y_reg_factor, void *stream);

/* LinearSmoothOp::apply (quoc/linearSmoothOp.cpp:591-612): one FD heat multigrid V-cycle with
 * tau = sigma^2/2 on each of nfields fields (in and out may alias). */
int msg_h1_smooth(int dtype, const void *in, void *out, int n, int nfields, double sigma, void *stream);

/* DeformImage (quoc/deformations.h:62-101, mode 0: constant `ext` outside, optional nearest
 * neighbour), FastDeformImage (L140-162, mode 1: clamp) and the NaN-extended warp (L226-256,
 * mode 2). phi_shared: one field for all nimg images. */
int msg_warp(int dtype, const void *img, const void *phi, void *out, int n, int nimg, int phi_shared, int mode,
             double ext, int nearest, void *stream);

/* CellCenteredRestrictOp / CellCenteredProlongOp (quoc/cellCenteredGrid.h:157-210) and
 * ScalarArray<QC_2D>::resampleFrom (quoc/scalarArray.h:308-331, 1052-1060). */
int msg_restrict(int dtype, const void *fine, void *coarse, int n, int nfields, void *stream);
int msg_prolong(int dtype, const void *coarse, void *fine, int nc, int nfields, void *stream);
int msg_resample(int dtype, const void *src, void *dst, int nc, int n, int nfields, void *stream);

/* ConcatAndSmoothlyExtendDeformations (quoc/deformations.h:499-520, with SmoothlyExtendImage
 * L364-404): result = phi1(x + phi2(x)) + phi2(x), harmonically extended where undefined.
 * `result` holds the PCG start value on entry, as the reference's output argument does. */
int msg_compose(int dtype, const void *phi1, const void *phi2, void *result, int n, void *stream);

/* Reconstruction statistics of SeriesMatching::averageSeries (matchSeries.h:404-444;
 * MultiVector::getMedianVecOverComponents, aol/multiVector.cpp:99-117): per pixel mean over
 * finite samples (+Inf marks a missing one), sample count and median; `fill` where no sample. */
int msg_series_stats(int dtype, const void *stack, int nimg, int n, double fill, void *mean, void *count,
                     void *median, void *stream);

/* Input normalisation of prepareMultilevelArray (registration.h:331-347): saturation clamp at the
 * (pct/200) quantiles and [0,1] scaling (pct >= 0), else shift-to-zero and divide by the maximum
 * (aol/vec.h:864-870, 1084-1089). In place on nimg device images. */
int msg_normalize(int dtype, void *img, int n, int nimg, double pct, int min_to_zero, void *stream);

/* ScalarArray<QC_2D>::resampleFrom (quoc/scalarArray.h:308-331) of a w x h source (w, h >= 2) onto an
 * n x n array: bilinear interpolation at x (w-1)/(n-1), y (h-1)/(n-1) ("resizeInput",
 * registration.h:369-372). Device arrays, x fastest. */
int msg_resample_rect(int dtype, const void *src, int w, int h, void *dst, int n, void *stream);

/* qc::approxInverseDeformation (quoc/deformations.h:341-355, TransformFunction<QC_2D>::transform,
 * quoc/auxiliary.cpp:33-176): the identity pushed forward through the displacement phi (two
 * triangles per element, the first covering triangle in element order wins), minus the identity;
 * uncovered nodes take the midpoint of the covered values' range. phi, inv: distinct device fields
 * (x component, then y) of n x n. */
int msg_approx_inverse(int dtype, const void *phi, void *inv, int n, void *stream);

/* PhaseCorrelationRegistration::registerImages (quoc/paramReg.h:2057-2096) over
 * Convolution::phaseCorrelation (quoc/convolution.h:1092-1114): the integer translation b (pixels,
 * shift[0] along x, shift[1] along y, each in (-n/2, n/2]) with reference(. + b) ~ templ, taken at
 * the peak of the phase correlation (the first maximum in x-fastest order); max_shift > 0 searches
 * |b|_inf <= max_shift (< n) on the plain cross-correlation instead. Device images of n x n,
 * n <= 4096; corr: NULL or n*n device doubles receiving the correlation surface. */
int msg_register_translation(int dtype, const void *reference, const void *templ, int n, int max_shift, int *shift,
                             double *corr, void *stream);

/* ------------------------------------------------------------------------------------------------
 * Series matching: one stage of SeriesMatching::matchAndAverageSeries (matchSeries.h:87-269,
 * 353-451, 565-570) on device-resident frames.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
  int start_level, stop_level, refine_start_level, refine_stop_level, precision_level;
  double lambda, lambda_factor;       /* lambda_factor <= 0: absent                          */
  int max_iter;
  double stop_eps;
  double saturation_pct;              /* "enhanceContrastSaturationPercentage"; < 0: absent   */
  int normalize_min_to_zero;          /* "normalizeMinToZero"                                 */
  int dont_normalize;                 /* "dontNormalizeInputImages"                           */
  int resample_prolong;               /* "resampleInsteadOfProlongateDeformation"             */
  int dont_accumulate;                /* "dontAccumulateDeformation"                          */
  int spec;                           /* Armijo candidates per energy pass (0: default)       */
  double pre_smooth_sigma;            /* "preSmoothSigma" (pixels); <= 0: none                */
  int pairwise_batch;                 /* pairwise solves per batch; 0: automatic. The chain   */
                                      /* starts on the first finished batch and overlaps the  */
                                      /* later ones (results do not depend on it)             */
  int data_term;                      /* MSG_SSD ("deformationModel 22") or MSG_NCC (0)       */
  int average_ext_mean;               /* "deformForAverageExtendedWithMean": outside the      */
                                      /* domain a frame takes its own mean instead of no sample*/
  int average_nearest;                /* "deformForAverageUsingNearestNeighborInterpolation"  */
  int reduce_deformations;            /* "reduceDeformations": after the matching, the        */
  int reduce_max_iter;                /* deformation reducing the consistency energy of all   */
  double reduce_stop_eps;             /* frames is found and applied first when averaging     */
  const void *pair0_init;             /* NULL or a device field: template 0's starting        */
                                      /* deformation (stage > 1 with reduceDeformations)      */
  int corr_init;                      /* "useCorrelationToInitTranslation": every pairwise    */
  int max_corr_shift;                 /* solve starts from the translation found by phase     */
                                      /* correlation of its two frames ("maxCorrShift" > 0:   */
                                      /* the plain correlation peak within that many pixels)  */
  int calc_inverse;                   /* "calcInverseDeformation": after frame i's (refined)  */
                                      /* deformation, a solve over the refine levels with the */
                                      /* two images' roles swapped, started from its          */
                                      /* approximate inverse (matchSeries.h:247-259)          */
  void *inverse_out;                  /* NULL or n device fields: those inverse solves' result*/
  int alt_start_level;                /* "altStartLevel" (0, or start_level: none): every     */
  int energy_as_alt_criterion;        /* pairwise solve is repeated from zero at that level;  */
                                      /* the first result is kept when it leaves fewer        */
                                      /* template pixels outside the domain (or, with         */
                                      /* "energyAsAltCriterion", has the lower energy)        */
                                      /* (matchSeries.h:158-203)                              */
  void *first_try_out;                /* NULL or n device fields: the start_level solves      */
  const void *raw_templates;          /* NULL, or n device images the reconstruction warps    */
                                      /* instead of `templates` (when the caller prepared the */
                                      /* matching inputs itself: normalised at their own      */
                                      /* resolution, then restricted, registration.h:369-399) */
  double y_reg_factor;                /* MSG_XSHIFT: "yRegFactor"                             */
  int keep_start_transformation;      /* the series loop sees, after each solve, the          */
                                      /* deformation the solve started from: matchSeismic-    */
                                      /* Series' dyadic path, whose registration object hands */
                                      /* back only what was set (getTransformation returns    */
                                      /* _phi, which the full-resolution step alone updates,  */
                                      /* matchSeries.h:1078-1096, 1177-1213). A pairing       */
                                      /* solved here composes as zero, a loaded one as loaded;*/
                                      /* the refinement's start, not its result, stays the    */
                                      /* accumulated deformation; each frame is averaged with */
                                      /* its refined deformation (the files it was saved to)  */
} msg_series_params;

/* Optional per-level records of msg_match_series: what the reference writes after every level of
 * every solve as <i>/ or <i>-r/deformation_<LL>_<c>.dat.bz2 and energy_<LL>.txt
 * (RegistrationConfigurator::writeCurrentRegistration via descentOnCurrentGrid,
 * registration.h:1092-1112; GradientDescentBase::writeEnergy, aol/gradientDescent.h:254-271).
 * Records are ordered by solve — the pairwise solves i = 0..npw-1 (npw = 1 with pairwise_in,
 * else n) over levels start..stop, then the refinements i = 1..n-1 over
 * refine_start..refine_stop (none with dont_accumulate), then with calc_inverse the inverse solves
 * i = 0..n-1 over refine_start..refine_stop, then with alt_start_level the repeated pairwise solves
 * i = 0..npw-1 over alt_start_level..stop — and by ascending level within a solve.
 * Every pointer may be NULL. All host memory. */
typedef struct {
  double *fields;         /* 2*N_l^2 doubles per record (x component, then y), back to back    */
  msg_iter_trace *iters;  /* max_iter records per record slot (the first `iterations` are set) */
  double *start;          /* 2 per record: E and lambda_l*E_reg at the start of the level     */
  int *iterations;        /* 1 per record                                                       */
  /* Called as each record completes (the reference writes a level's files when it finishes,
   * registration.h:1110-1111), with the level's field (2*N_l^2 doubles), its iteration records, the
   * count and the 2 start values; the pointers are valid during the call only. Called from the
   * library's threads (the pairwise solves and the chain run on different ones): it must be
   * thread-safe and should return quickly. NULL: none. */
  void (*on_level)(void *user, long long record, int level, const double *field, const msg_iter_trace *iters,
                   int iterations, const double *start);
  void *user;
} msg_series_levels;

/* templates: n device images (dtype) at precision_level; reference: one device image.
 * pairwise_in: NULL or n device fields; entry i >= 1 replaces template i's pairwise solve
 * ("reuseStage1Results"). Outputs (device, dtype): deformations (n fields, those used for the
 * average), pairwise_out (n fields or NULL), mean/count/median (one image each); energies: host,
 * n values; levels: NULL or the per-level records above; reduced: NULL or one device field, the
 * reduced deformation (reduce_deformations). */
int msg_match_series(const msg_series_params *p, int dtype, const void *templates, int n, const void *reference,
                     const void *pairwise_in, void *deformations, void *pairwise_out, double *energies,
                     void *mean, void *count, void *median, const msg_series_levels *levels, void *reduced);
/* Totals of the solves of the last msg_match_series in the process (no counterpart in the
 * reference): gradient-flow iterations, reference-equivalent energy evaluations and energy passes
 * over every pairwise, refinement and inverse solve, as msg_last_counts reports them per context;
 * node_iterations and node_evals weight each iteration and evaluation by its level's node count
 * (4^level), the unit of the path's byte model. Any output may be NULL. */
int msg_series_last_counts(long long *iterations, long long *evals, long long *passes, double *node_iterations,
                           double *node_evals);

/* Several independent series of possibly different frame sizes in one call (no counterpart in the
 * reference, whose program matches one series per run). Each job is one msg_match_series call's
 * inputs and outputs (device arrays of `dtype`, energies on the host); each series' results are
 * bitwise those of msg_match_series on it alone. Series with the same frame size and solve
 * parameters share contexts: their pairwise solves are batched together and their accumulate-and-
 * refine chains advance in lockstep, refinement i of every such series in one batched solve; groups
 * of different sizes run concurrently on their own streams. Not taken here (msg_match_series has
 * them): reduce_deformations, corr_init, calc_inverse, alt_start_level, pair0_init, raw_templates,
 * keep_start_transformation, MSG_XSHIFT, per-level records, reuse of earlier pairwise results. */
typedef struct {
  const msg_series_params *params;  /* the series' parameters (precision_level: its frame size) */
  const void *templates;            /* n images, 2^L x 2^L                                     */
  int n;
  const void *reference;            /* one image                                               */
  void *deformations;               /* n fields (output)                                       */
  double *energies;                 /* n host values (output)                                  */
  void *mean, *count, *median;       /* one image each (output)                                 */
} msg_series_job;
int msg_match_series_batch(int dtype, int njobs, const msg_series_job *jobs);
/* SeriesMatching::reduceDeformations (matchSeries.h:678-704): from zero, the displacement phi
 * minimising sum_i 1/2 int |psi_i(x + phi(x)) + phi(x)|^2 (ConsistencyEnergyOp and its variation,
 * AmbrosioTortorelli.h:687-800) by H1GradientDescent with filter width sqrt(10 h), nonlinear CG and
 * gradient-norm stopping (stop_eps; the reference's default 5e-7). deformations: n device fields at
 * level `level`; phi: device output field. energy/iterations: host outputs (may be NULL). */
int msg_reduce_deformations(int dtype, const void *deformations, int n, int level, int max_iter, double stop_eps,
                            void *phi, double *energy, int *iterations);
/* msg_match_series keeps its device contexts for reuse by later calls of the same shape (no
 * counterpart in the reference); this frees them. */
int msg_release_cache(void);

/* ------------------------------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference): a BF16 U-Net on MFMA that predicts a
 * deformation from a (template, reference) pair; the gradient flow then refines from it
 * (msg_set_deformations_device / msg_solve_level). Forward, backward and Adam run on the GPU; FP32
 * master weights, BF16 activations and weights in the products, FP32 accumulation.
 * Network: `scales` scales of two 3x3 convolutions + ReLU each (base_channels * 2^s channels), 2x2
 * average pooling down, nearest-neighbour upsampling + concatenation with the skip tensor up, a 1x1
 * head to 2 channels. Inputs: device FP32 [batch][2][H][W] (template, reference images); outputs and
 * targets: device FP32 [batch][2][H][W] displacements in pixels (x, then y). H and W must be multiples
 * of 32 * 2^(scales-1). Parameters: one flat FP32 vector, per convolution its weights
 * [9][Cout][Cin] (tap = 3 (dy + 1) + dx + 1; the input's 2 channels padded to 16 with zero weights)
 * and its biases, in network order, then the head's [2][C] weights and 2 biases.
 * ---------------------------------------------------------------------------------------------- */
typedef struct msg_unet msg_unet;
/* He-normal weights from `seed`, zero biases. scales 1..6; base_channels a multiple of 32, at most 256. */
int msg_unet_create(msg_unet **net, int scales, int base_channels, unsigned long long seed);
int msg_unet_destroy(msg_unet *net);
int msg_unet_param_count(msg_unet *net, long long *count);
/* Host copies of the flat parameter vector (setting also resets the Adam state). */
int msg_unet_get_params(msg_unet *net, float *host);
int msg_unet_set_params(msg_unet *net, const float *host);
/* out = net(in). `stream` (or NULL) is synchronised before the network reads `in`; the call returns
 * when `out` is written. */
int msg_unet_forward(msg_unet *net, const float *in, float *out, int batch, int H, int W, void *stream);
/* Gradient of the loss 1/(2 n) sum (net(in) - target)^2 (n = 2 batch H W) with respect to every
 * parameter, kept for msg_unet_adam_step; loss and grad_host (param_count floats) may be NULL. */
int msg_unet_gradient(msg_unet *net, const float *in, const float *target, int batch, int H, int W, double *loss,
                      float *grad_host, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr on the last gradient. */
int msg_unet_adam_step(msg_unet *net, float lr);

/* ------------------------------------------------------------------------------------------------
 * Kernel timing (no counterpart in the reference; used by the benchmark). With class bit k set in
 * `mask`, every launch of kernel class k is bracketed by HIP events on its own stream. Classes:
 * 0 energy, 1 variation, 2 H1 V-cycle, 3 (none: the CG direction is formed inside the energy and
 * update kernels), 4 update, 5 control, 6 composition.
 * msg_profile_read returns the summed event time, launch count and algorithmic bytes of a class.
 * ---------------------------------------------------------------------------------------------- */
int msg_profile_enable(unsigned mask);
int msg_profile_reset(void);
int msg_profile_read(int cls, double *ms, long long *launches, double *bytes);
/* Times the energy pass of a context's level solve (no counterpart in the reference; used by the
 * benchmark's roofline): every pair evaluates `ncand` (1..3) step sizes (1, 1/2, ...) along the direction the
 * last solve at this level left in the context, `reps` launches back to back on the context's stream
 * between two HIP events. Outputs the mean launch time (ms) and the algorithmic bytes of one launch
 * (reference, template, deformation and direction: 6 fields per pair). Leaves the pair state unusable
 * for resuming; the next solve starts afresh as always. */
int msg_time_energy_pass(msg_ctx *ctx, int level, int ncand, int reps, double *ms_per_launch,
                         double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
