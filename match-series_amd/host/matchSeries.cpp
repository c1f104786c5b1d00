// matchSeries <file.par> [ACTION] — the reference's command line (projects/electronMicroscopy/
// matchSeries.cpp:11-209, matchSeries.h:1249-1320) over the gfx950 library's C ABI.
//
// deformationModel 22 (SSD data term) and 0 (normalised cross-correlation), each with the Dirichlet
// regulariser and the H1 gradient flow, are computed on the GPU. ACTION 0 (match and average), 1 (only
// average), 2 (analyse the saved deformations) and 3 (apply their sum to the templates) as the
// reference's SeriesMatching::doAction (matchSeries.h:782-799). The output tree is the
// reference's: <saveDirectory>/{log.txt, parameter-dump.txt, stage<k>/...}. Errors are reported as
// the reference does (message on stderr, exit code 0); MSG_STRICT=1 makes them exit with 1.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <streambuf>
#include <string>
#include <vector>

#include "array_io.h"
#include "image_formats.h"
#include "ms_gpu.h"
#include "params.h"

using namespace mshost;

namespace {

// matchSeismicSeries (built from this source with MS_SEISMIC): the same series loop over the x-shift
// registration of XShiftNonDyadicRegistrationMultilevelDescent (matchSeismicSeries.cpp:10-86).
#ifdef MS_SEISMIC
constexpr bool kSeismic = true;
#else
constexpr bool kSeismic = false;
#endif

// Duplicates everything written to a stream into log.txt (aol::AdditionalOutputToFile,
// multiStreambuf.h:49-63).
class TeeBuf : public std::streambuf {
 public:
  TeeBuf(std::streambuf *a, std::streambuf *b) : a_(a), b_(b) {}
  int overflow(int c) override {
    if (c == EOF) return !EOF;
    const int r1 = a_->sputc((char)c), r2 = b_->sputc((char)c);
    return (r1 == EOF || r2 == EOF) ? EOF : c;
  }
  int sync() override { return (a_->pubsync() == 0 && b_->pubsync() == 0) ? 0 : -1; }

 private:
  std::streambuf *a_, *b_;
};

void hipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess) MSH_THROW(strprintf("%s: %s", what, hipGetErrorString(e)));
}
void msgCheck(int rc) {
  if (rc != MSG_OK) MSH_THROW(strprintf("libmsgpu: %s", msg_last_error()));
}

struct DeviceBuffer {
  void *p = nullptr;
  explicit DeviceBuffer(size_t bytes) { hipCheck(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc"); }
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};

// host float64 -> device field type
void upload(void *dst, const double *src, size_t n, int dtype) {
  if (dtype == MSG_F64) {
    hipCheck(hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice), "upload");
  } else {
    std::vector<float> t(src, src + n);
    hipCheck(hipMemcpy(dst, t.data(), n * 4, hipMemcpyHostToDevice), "upload");
  }
}
void download(double *dst, const void *src, size_t n, int dtype) {
  if (dtype == MSG_F64) {
    hipCheck(hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost), "download");
  } else {
    std::vector<float> t(n);
    hipCheck(hipMemcpy(t.data(), src, n * 4, hipMemcpyDeviceToHost), "download");
    for (size_t i = 0; i < n; ++i) dst[i] = t[i];
  }
}

std::string dirOf(const ParameterParser &p) { return p.getStringExpandTilde("saveDirectory"); }

// Options of the reference that change the computation but are not implemented here.
void rejectUnsupported(const ParameterParser &p) {
  const char *flags[] = {"reverseRolesInSeriesMatching", "calcIQ", "calcPrecision",
                         "performNoiseAnalysis"};
  for (const char *f : flags)
    if (p.checkAndGetBool(f)) MSH_THROW(strprintf("\"%s 1\" is not supported by this build", f));
  // "initial-deformation": the registration object loads it (registration.h:495-500), and the series
  // loop starts every pairing from zero regardless (matchSeries.h:111, 137): only the loading shows
  if (p.hasVariable("initial-deformation"))
    for (int c = 0; c < 2; ++c) (void)loadImage(p.getString("initial-deformation", c));
  // the reference itself does not implement a positive threshold (registration.h:604-610)
  if (p.getDoubleOrDefault("averageXDerivNormThreshold", 0) > 0) MSH_THROW("averageXDerivNormThreshold: not implemented");
  if (p.hasVariable("minimizationAlgo") && p.getString("minimizationAlgo") != "GradientDescent")
    MSH_THROW("only \"minimizationAlgo GradientDescent\" is supported by this build");
}

bool useAltStartLevel(const ParameterParser &p) {  // SeriesMatching::_useAltStartLevel (matchSeries.h:45)
  return p.hasVariable("altStartLevel") && p.getInt("altStartLevel") != p.getInt("startLevel");
}

msg_series_params seriesParams(const ParameterParser &p) {
  msg_series_params s;
  memset(&s, 0, sizeof(s));
  s.precision_level = p.getInt("precisionLevel");
  s.start_level = p.getInt("startLevel");
  s.stop_level = p.getInt("stopLevel");
  s.refine_start_level = p.getIntOrDefault("refineStartLevel", s.precision_level);
  s.refine_stop_level = p.getIntOrDefault("refineStopLevel", s.precision_level);
  s.lambda = p.getDouble("lambda");
  s.lambda_factor = p.hasVariable("lambdaFactor") ? p.getDouble("lambdaFactor") : -1.0;
  s.max_iter = p.getIntOrDefault("maxGDIterations", 1000);  // StandardRegistration defaults (registration.h:940-950)
  s.stop_eps = p.getDoubleOrDefault("stopEpsilon", 0.0);
  s.saturation_pct = p.hasVariable("enhanceContrastSaturationPercentage")
                         ? p.getDouble("enhanceContrastSaturationPercentage") : -1.0;
  s.normalize_min_to_zero = p.checkAndGetBool("normalizeMinToZero");
  s.dont_normalize = p.checkAndGetBool("dontNormalizeInputImages");
  s.resample_prolong = p.checkAndGetBool("resampleInsteadOfProlongateDeformation");
  s.dont_accumulate = p.checkAndGetBool("dontAccumulateDeformation");
  s.spec = p.getIntOrDefault("gpuArmijoCandidates", 0);
  s.pre_smooth_sigma = p.getDoubleOrDefault("preSmoothSigma", 0.0);
  s.pairwise_batch = p.getIntOrDefault("gpuPairwiseBatch", 0);
  s.data_term = p.getIntOrDefault("deformationModel", 0) == 22 ? MSG_SSD : MSG_NCC;
  s.average_ext_mean = p.checkAndGetBool("deformForAverageExtendedWithMean");
  s.average_nearest = p.checkAndGetBool("deformForAverageUsingNearestNeighborInterpolation");
  s.reduce_deformations = p.checkAndGetBool("reduceDeformations");
  s.reduce_max_iter = p.getIntOrDefault("maxGDIterations", 1000);  // SeriesMatching::reduceDeformations
  s.reduce_stop_eps = p.getDoubleOrDefault("stopEpsilon", 5e-7);   // (matchSeries.h:697)
  s.pair0_init = nullptr;
  // every new pairing starts from the phase-correlation translation of its frames (matchSeries.h:145-153)
  s.corr_init = p.checkAndGetBool("useCorrelationToInitTranslation");
  s.max_corr_shift = s.corr_init ? p.getInt("maxCorrShift") : 0;
  // after each frame, the inverse problem solved from the approximate inverse (matchSeries.h:247-259)
  s.calc_inverse = p.checkAndGetBool("calcInverseDeformation");
  s.inverse_out = nullptr;
  // a second pairwise solve from altStartLevel, the better one kept (matchSeries.h:158-203)
  s.alt_start_level = useAltStartLevel(p) ? p.getInt("altStartLevel") : 0;
  s.energy_as_alt_criterion = p.checkAndGetBool("energyAsAltCriterion");
  s.first_try_out = nullptr;
  if (kSeismic) {
    // XShiftRegistrationMultilevelDescent::descentOnCurrentGrid (matchSeries.h:953-972): findXShift with
    // "lambda" on every level, "yRegFactor", maxGDIterations (default 500), stopEpsilon 0; the frames
    // scaled to [0, 1] (NonDyadicRegistrationMultilevelDescent::loadAndPrepareImage, matchSeries.h:1040-1053)
    s.data_term = MSG_XSHIFT;
    s.y_reg_factor = p.getDouble("yRegFactor");
    s.lambda_factor = -1.0;
    s.max_iter = p.getIntOrDefault("maxGDIterations", 500);
    s.stop_eps = 0.0;
    s.saturation_pct = -1.0;
    s.normalize_min_to_zero = 1;
    s.pre_smooth_sigma = 0;
    s.keep_start_transformation = 1;
  }
  return s;
}

// Field storage: FP32 by default; matchSeismicSeries defaults to FP64 (its row kernels compute in
// FP64 whatever the storage, so FP32 fields save little there and move its short solves).
int fieldType(const ParameterParser &p) {
  return p.getIntOrDefault("gpuFieldPrecision", kSeismic ? 64 : 32) == 64 ? MSG_F64 : MSG_F32;
}

// An input image as RegistrationMultilevelDescentInterface::initMultilevelArrayFromImage prepares it
// (registration.h:330-399): resampled onto the grid of level max(L, log2 width) with resizeInput, cut
// out of the input at (cropStartX, cropStartY) with cropInput, else taken at its own dyadic level
// k >= L; normalised there (scale: prepareMultilevelArray's
// saturation clamp or min/max scaling), pre-smoothed there (smooth: preSmoothSigma in pixels of
// that level), then restricted to level L. Writes 2^L x 2^L values of the field type to dst.
void prepareFrame(const ParameterParser &p, const Image &im, const std::string &name, int L, int dt, bool scale,
                  bool smooth, bool noCrop, void *dst) {
  const int N = 1 << L, es = dt == MSG_F64 ? 8 : 4;
  int k = L;
  std::vector<double> full;
  std::unique_ptr<DeviceBuffer> cur;
  if (!noCrop && p.checkAndGetBool("resizeInput")) {
    int inLevel = 0;
    while ((2 << inLevel) <= im.nx) ++inLevel;  // qc::logBaseTwo
    k = std::max(L, inLevel);
    const int n = 1 << k;
    if (im.nx < 2 || im.ny < 2) MSH_THROW(strprintf("resizeInput: %s is too small", name.c_str()));
    DeviceBuffer src((size_t)es * im.nx * im.ny);
    upload(src.p, im.data.data(), (size_t)im.nx * im.ny, dt);
    cur.reset(new DeviceBuffer((size_t)es * n * n));
    msgCheck(msg_resample_rect(dt, src.p, im.nx, im.ny, cur->p, n, nullptr));
  } else if (!noCrop && p.checkAndGetBool("cropInput")) {
    const int sx = p.getInt("cropStartX"), sy = p.getInt("cropStartY");
    const int w = std::min(im.nx, N), h = std::min(im.ny, N);
    if (sx < 0 || sy < 0 || sx + w > im.nx || sy + h > im.ny)
      MSH_THROW(strprintf("cropInput: the %dx%d block at (%d, %d) is not inside %s (%dx%d)", w, h, sx, sy,
                          name.c_str(), im.nx, im.ny));
    full.assign((size_t)N * N, 0.0);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) full[(size_t)y * N + x] = im.data[(size_t)(sy + y) * im.nx + sx + x];
  } else {
    k = 0;
    while ((2 << k) <= im.nx) ++k;  // qc::logBaseTwo
    if (im.nx != (1 << k) || im.ny != (1 << k))
      MSH_THROW(strprintf("Input data size (%dx%d of %s) doesn't match the size of the computational grid. Either "
                          "use a different type of grid or change the size with the input data with \"resizeInput\" "
                          "or \"cropInput\".", im.nx, im.ny, name.c_str()));
    if (k < L)
      MSH_THROW("Input data size too small. Note: Unless \"resizeInput\" is used, \"precisionLevel\" may not be bigger "
                "than the grid level of the input data.");
    full = im.data;
  }
  const int n = 1 << k;
  if (!cur) {
    cur.reset(new DeviceBuffer((size_t)es * n * n));
    upload(cur->p, full.data(), (size_t)n * n, dt);
  }
  if (scale)
    msgCheck(msg_normalize(dt, cur->p, n, 1,
                           p.hasVariable("enhanceContrastSaturationPercentage")
                               ? p.getDouble("enhanceContrastSaturationPercentage") : -1.0,
                           p.checkAndGetBool("normalizeMinToZero"), nullptr));
  const double sigma = p.getDoubleOrDefault("preSmoothSigma", 0.0);
  if (smooth && sigma > 0) msgCheck(msg_h1_smooth(dt, cur->p, cur->p, n, 1, sigma / (double)(n - 1), nullptr));
  for (int m = n; m > N; m /= 2) {
    std::unique_ptr<DeviceBuffer> half(new DeviceBuffer((size_t)es * (m / 2) * (m / 2)));
    msgCheck(msg_restrict(dt, cur->p, half->p, m, 1, nullptr));
    cur.swap(half);
  }
  hipCheck(hipMemcpy(dst, cur->p, (size_t)es * N * N, hipMemcpyDeviceToDevice), "copy");
}

// The raw frame at level L (loadAndPrepareImage(..., NoScaling, ..., NoSmoothing), as averageSeries
// reads the frames) into a device buffer of the field type.
void loadRaw(const ParameterParser &p, const std::string &f, int L, int dt, void *dst, bool noCrop = false) {
  prepareFrame(p, loadImage(f), f, L, dt, false, false, noCrop, dst);
}

// <dir><i>[-r]/deformation_<LL> (SeriesMatching::createDeformationBaseFileName, matchSeries.h:58-67)
std::string deformationBase(const ParameterParser &p, const std::string &dir, int i, int maxDepth) {
  if (useAltStartLevel(p) && i == 0) return dir + "deformation-000";
  const bool refined = !p.checkAndGetBool("dontAccumulateDeformation") && i > 0;
  return strprintf("%s%d%s/deformation_%02d", dir.c_str(), i, refined ? "-r" : "",
                   refined ? maxDepth : p.getInt("stopLevel"));
}

void writeReconstruction(const std::string &dir, const std::vector<double> &mean, const std::vector<double> &median,
                         const std::vector<double> &count, int N) {
  saveP9(dir + "average.q2bz", mean.data(), N, N);
  saveP9(dir + "median.q2bz", median.data(), N, N);
  saveP9(dir + "numSamples.q2bz", count.data(), N, N);
  savePNGScaled(dir + "average.png", mean.data(), N, N);
  savePNGScaled(dir + "median.png", median.data(), N, N);
  savePNGScaled(dir + "numSamples.png", count.data(), N, N);
}

// aol::getBaseFileName: the file name without directory and last extension
std::string baseFileName(const std::string &f) {
  const size_t slash = f.rfind('/');
  std::string b = slash == std::string::npos ? f : f.substr(slash + 1);
  const size_t dot = b.rfind('.');
  return dot == std::string::npos ? b : b.substr(0, dot);
}

double meanOf(const std::vector<double> &v) {
  double s = 0;
  for (double x : v) s += x;
  return v.empty() ? 0.0 : s / v.size();  // ScalarArray::getMeanValue
}

// The deformed templates SeriesMatching::averageSeries averages (matchSeries.h:381-392): with
// reduceDeformations the reduced deformation first (+Inf outside), then each template's own
// deformation, +Inf outside (or the image's mean with deformForAverageExtendedWithMean), optionally
// nearest-neighbour. img: the n raw templates on the device (the reduced pre-warp replaces them).
void deformForAverage(const ParameterParser &p, int n, int N, int dt, void *img, const void *defs,
                      const void *reduced, void *stack) {
  const size_t nn = (size_t)N * N;
  const int es = dt == MSG_F64 ? 8 : 4;
  const bool extMean = p.checkAndGetBool("deformForAverageExtendedWithMean");
  const int nearest = p.checkAndGetBool("deformForAverageUsingNearestNeighborInterpolation") ? 1 : 0;
  if (reduced) {
    DeviceBuffer pre(es * nn * n);
    msgCheck(msg_warp(dt, img, reduced, pre.p, N, n, 1, 0, INFINITY, 0, nullptr));
    hipCheck(hipMemcpy(img, pre.p, es * nn * n, hipMemcpyDeviceToDevice), "copy");
  }
  if (extMean) {
    for (int i = 0; i < n; ++i) {
      std::vector<double> h(nn);
      download(h.data(), (char *)img + (size_t)i * es * nn, nn, dt);
      msgCheck(msg_warp(dt, (char *)img + (size_t)i * es * nn, (const char *)defs + (size_t)i * es * 2 * nn,
                        (char *)stack + (size_t)i * es * nn, N, 1, 0, 0, meanOf(h), nearest, nullptr));
    }
  } else {
    msgCheck(msg_warp(dt, img, defs, stack, N, n, 0, 0, INFINITY, nearest, nullptr));
  }
}

// mean / count / median of the first k deformed templates, `fill` where no sample, written with
// the iteration suffix "_<k>" (three digits) unless iter < 0 (saveAverageMedianAndNumSamples,
// matchSeries.h:271-279).
void reconstructAndWrite(const std::string &dir, const void *stack, int k, int N, int dt, double fill, int iter) {
  const size_t nn = (size_t)N * N;
  const int es = dt == MSG_F64 ? 8 : 4;
  DeviceBuffer mean(es * nn), cnt(es * nn), med(es * nn);
  msgCheck(msg_series_stats(dt, stack, k, N, fill, mean.p, cnt.p, med.p, nullptr));
  std::vector<double> hm(nn), hc(nn), hmed(nn);
  download(hm.data(), mean.p, nn, dt);
  download(hc.data(), cnt.p, nn, dt);
  download(hmed.data(), med.p, nn, dt);
  const std::string sfx = iter < 0 ? std::string() : strprintf("_%03d", iter);
  for (int f = 0; f < 3; ++f) {
    const char *nm = f == 0 ? "average" : f == 1 ? "median" : "numSamples";
    const std::vector<double> &a = f == 0 ? hm : f == 1 ? hmed : hc;
    saveP9(dir + nm + sfx + ".q2bz", a.data(), N, N);
    savePNGScaled(dir + nm + sfx + ".png", a.data(), N, N);
  }
}

// saveNamedDeformedTemplate / saveNamedDeformedTemplateHelper (matchSeries.h:332-351): DM3/DM4 inputs
// optionally back into a copy of their own container, everything else as <basename>.q2bz in the
// save directory.
void saveNamedDeformedTemplate(const ParameterParser &p, const std::string &file, const double *a, int N) {
  const std::string dir = dirOf(p);
  const bool dm = file.size() >= 3 && (file.compare(file.size() - 3, 3, "dm3") == 0 ||
                                       file.compare(file.size() - 3, 3, "dm4") == 0);
  if (dm && p.checkAndGetBool("saveNamedDeformedDMXTemplatesAsDMX"))
    saveInDMContainer(file, a, N, N, dir + baseFileName(file));
  else
    saveP9(dir + baseFileName(file) + ".q2bz", a, N, N);
}

// The optional outputs of SeriesMatching::averageSeries (matchSeries.h:365-423) besides the final
// reconstruction: <basename>.q2bz, each template deformed by its own deformation with 0 (or its mean)
// outside (saveNamedDeformedTemplates; in stage 1 also the undeformed reference);
// <basename>_def_<c>.dat.bz2, its deformation (saveNamedDeformations); the reconstruction of the
// first k templates after every averageSaveIncrement of them. img: raw templates (device), defs:
// their deformations (device), reduced: NULL or the reduced deformation (device).
void seriesExtras(const ParameterParser &p, int stage, const std::vector<std::string> &files, const void *img,
                  const void *defs, const void *reduced, int N, int dt) {
  const std::string dir = dirOf(p);
  const int n = (int)files.size();
  const size_t nn = (size_t)N * N;
  const int es = dt == MSG_F64 ? 8 : 4;
  const bool namedTpl = p.checkAndGetBool("saveNamedDeformedTemplates");
  const bool namedDef = p.checkAndGetBool("saveNamedDeformations");
  const int inc = p.hasVariable("averageSaveIncrement") ? p.getInt("averageSaveIncrement") : n;
  if (inc <= 0) MSH_THROW("'averageSaveIncrement' must be bigger than zero.");
  auto saveNamed = [&](const std::string &file, const double *a) { saveNamedDeformedTemplate(p, file, a, N); };
  if (namedTpl && stage == 1) {
    const std::string ref = p.getStringExpandTilde("reference");
    int L = 0;
    while ((2 << L) <= N) ++L;
    DeviceBuffer r(es * nn);
    loadRaw(p, ref, L, dt, r.p);
    std::vector<double> h(nn);
    download(h.data(), r.p, nn, dt);
    saveNamed(ref, h.data());
  }
  if (namedTpl) {
    const bool extMean = p.checkAndGetBool("saveNamedDeformedTemplatesExtendedWithMean");
    const int nearest = p.checkAndGetBool("saveNamedDeformedTemplatesUsingNearestNeighborInterpolation") ? 1 : 0;
    DeviceBuffer out(es * nn);
    std::vector<double> h(nn);
    for (int i = 0; i < n; ++i) {
      const char *im = (const char *)img + (size_t)i * es * nn;
      double ext = 0;
      if (extMean) {
        download(h.data(), im, nn, dt);
        ext = meanOf(h);
      }
      msgCheck(msg_warp(dt, im, (const char *)defs + (size_t)i * es * 2 * nn, out.p, N, 1, 0, 0, ext, nearest, nullptr));
      download(h.data(), out.p, nn, dt);
      saveNamed(files[i], h.data());
    }
  }
  if (namedDef) {
    std::vector<double> h(2 * nn);
    for (int i = 0; i < n; ++i) {
      download(h.data(), (const char *)defs + (size_t)i * es * 2 * nn, 2 * nn, dt);
      saveField(dir + baseFileName(files[i]) + "_def_%d.dat.bz2", h.data(), N);
    }
  }
  if (inc < n) {
    DeviceBuffer work(es * nn * n), stack(es * nn * n);
    hipCheck(hipMemcpy(work.p, img, es * nn * n, hipMemcpyDeviceToDevice), "copy");
    deformForAverage(p, n, N, dt, work.p, defs, reduced, stack.p);
    std::vector<double> h(nn);
    for (int k = inc; k < n; k += inc) {
      download(h.data(), (const char *)img + (size_t)(k - 1) * es * nn, nn, dt);
      reconstructAndWrite(dir, stack.p, k, N, dt, meanOf(h), k);
    }
  }
}

// ACTION 1: SeriesMatching::averageSeries (matchSeries.h:353-451) from saved deformations.
void averageOnly(const ParameterParser &p, int stage) {
  const std::string dir = dirOf(p);
  const int L = p.getInt("precisionLevel"), N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = fieldType(p), es = dt == MSG_F64 ? 8 : 4;
  const std::vector<std::string> files = createTemplateFileNameList(p);
  const int n = (int)files.size();
  DeviceBuffer img(es * nn * n), raw(es * nn * n), defs(es * 2 * nn * n), stack(es * nn * n);
  double lastMean = 0;
  for (int i = 0; i < n; ++i) {
    loadRaw(p, files[i], L, dt, (char *)img.p + (size_t)i * es * nn);
    std::vector<double> phi;
    loadField(deformationBase(p, dir, i, L) + "_%d.dat.bz2", N, phi);
    upload((char *)defs.p + (size_t)i * es * 2 * nn, phi.data(), 2 * nn, dt);
    if (i == n - 1) {
      std::vector<double> h(nn);
      download(h.data(), (char *)img.p + (size_t)i * es * nn, nn, dt);
      lastMean = meanOf(h);
    }
  }
  hipCheck(hipMemcpy(raw.p, img.p, es * nn * n, hipMemcpyDeviceToDevice), "copy");
  std::unique_ptr<DeviceBuffer> redD;
  if (p.checkAndGetBool("reduceDeformations")) {  // the reduced deformation first (matchSeries.h:381-392)
    std::vector<double> red;
    loadField(dir + "reduceDef_%d.dat.bz2", N, red);
    redD.reset(new DeviceBuffer(es * 2 * nn));
    upload(redD->p, red.data(), 2 * nn, dt);
  }
  deformForAverage(p, n, N, dt, img.p, defs.p, redD ? redD->p : nullptr, stack.p);
  seriesExtras(p, stage, files, raw.p, defs.p, redD ? redD->p : nullptr, N, dt);
  reconstructAndWrite(dir, stack.p, n, N, dt, lastMean, -1);
}

// ACTION 0 of one stage: SeriesMatching::matchAndAverageSeries (matchSeries.h:565-570).
// Writes the per-level files of the solves while the GPU runs: the library hands each finished level
// to onLevel (from its own threads), which queues the file writing (bzip2 of the field) for a few
// worker threads. Holding nothing beyond the queue, the stage needs no host copy of all its levels.
class LevelWriter {
 public:
  struct Slot {
    std::string sub;  // directory of the solve ("<i>/", "<i>-r/", ...)
    int level;
  };
  LevelWriter(std::vector<Slot> slots, const msg_series_params &sp, int nthreads)
      : slots_(std::move(slots)), sp_(sp) {
    for (int k = 0; k < nthreads; ++k) threads_.emplace_back([this] { loop(); });
  }
  ~LevelWriter() { join(); }
  // Waits until every queued file is written; rethrows the first failure.
  void finish() {
    join();
    if (!err_.empty()) MSH_THROW(err_);
  }
  static void onLevel(void *user, long long record, int level, const double *field, const msg_iter_trace *iters,
                      int iterations, const double *start) {
    auto *w = static_cast<LevelWriter *>(user);
    const int m = 1 << level;
    auto job = std::make_shared<Job>();
    job->record = record;
    job->field.assign(field, field + (size_t)2 * m * m);
    if (iters) job->iters.assign(iters, iters + std::max(iterations, 0));
    job->start[0] = start[0];
    job->start[1] = start[1];
    std::lock_guard<std::mutex> g(w->mu_);
    w->q_.push_back(job);
    w->cv_.notify_one();
  }

 private:
  void join() {
    {
      std::unique_lock<std::mutex> l(mu_);
      done_ = true;
      cv_.notify_all();
    }
    for (auto &t : threads_)
      if (t.joinable()) t.join();
    threads_.clear();
  }
  struct Job {
    long long record;
    std::vector<double> field;
    std::vector<msg_iter_trace> iters;
    double start[2];
  };
  void loop() {
    for (;;) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [this] { return done_ || !q_.empty(); });
        if (q_.empty()) return;
        j = q_.front();
        q_.pop_front();
      }
      try {
        write(*j);
      } catch (std::exception &e) {
        std::lock_guard<std::mutex> g(mu_);
        if (err_.empty()) err_ = e.what();
      }
    }
  }
  // deformation_<LL>_<c>.dat.bz2 and energy_<LL>.txt as GradientDescentBase::writeEnergy prints them:
  // "k lambda*E_reg E" from k = -1; NCC (a GenEnergyOp too) puts its last value first, E - lambda*E_reg
  void write(const Job &j) const {
    if (j.record < 0 || j.record >= (long long)slots_.size()) return;
    const Slot &sl = slots_[j.record];
    const int l = sl.level, m = 1 << l;
    saveField(sl.sub + strprintf("deformation_%02d_%%d.dat.bz2", l), j.field.data(), m);
    const double lam =
        sp_.lambda_factor > 0 ? sp_.lambda * std::pow(sp_.lambda_factor, sp_.stop_level - l) : sp_.lambda;
    std::ofstream e(sl.sub + strprintf("energy_%02d.txt", l));
    auto line = [&](int k, double regw, double E) {
      e << k << " ";
      if (sp_.data_term == MSG_NCC || sp_.data_term == MSG_XSHIFT) e << E - regw << " ";
      e << regw << " " << E << "\n";
    };
    line(-1, j.start[1], j.start[0]);
    for (size_t k = 0; k < j.iters.size(); ++k) line((int)k, lam * j.iters[k].ereg, j.iters[k].energy);
  }

  std::vector<Slot> slots_;
  msg_series_params sp_;
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<Job>> q_;
  bool done_ = false;
  std::string err_;
};

void matchAndAverage(const ParameterParser &p, bool loadStageOneResults, int stage) {
  const std::string dir = dirOf(p);
  const int L = p.getInt("precisionLevel"), N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = fieldType(p), es = dt == MSG_F64 ? 8 : 4;
  const std::vector<std::string> files = createTemplateFileNameList(p);
  const int n = (int)files.size();
  if (n == 0) MSH_THROW("no template images");
  msg_series_params sp = seriesParams(p);

  // reduceDeformations in a later stage: template 0 starts from the previous stage's reduced
  // deformation, from stage 3 on composed with the previous stage's deformation of template 0
  // (matchSeries.h:123-135)
  std::unique_ptr<DeviceBuffer> init0;
  if (sp.reduce_deformations && stage > 1) {
    std::vector<double> red;
    loadField(strprintf("%s../stage%d/reduceDef_%%d.dat.bz2", dir.c_str(), stage - 1), N, red);
    init0.reset(new DeviceBuffer(es * 2 * nn));
    upload(init0->p, red.data(), 2 * nn, dt);
    if (stage > 2) {
      std::vector<double> last;
      loadField(strprintf("%s../stage%d/0/deformation_%02d_%%d.dat.bz2", dir.c_str(), stage - 1, L), N, last);
      DeviceBuffer lastD(es * 2 * nn), redD(es * 2 * nn);
      upload(lastD.p, last.data(), 2 * nn, dt);
      upload(redD.p, red.data(), 2 * nn, dt);
      msgCheck(msg_compose(dt, lastD.p, redD.p, init0->p, N, nullptr));  // start value: the reduced deformation
    }
    sp.pair0_init = init0->p;
  }

  DeviceBuffer tpl(es * nn * n), ref(es * nn), defs(es * 2 * nn * n), pw(es * 2 * nn * n), mean(es * nn),
      cnt(es * nn), med(es * nn);
  // Frames of the grid's size are normalised inside the library. Larger dyadic frames and cropped
  // ones are prepared here as the reference prepares them (normalised at their own resolution, then
  // restricted); the reconstruction then reads the raw frames restricted alike.
  std::vector<Image> ims;
  for (int i = 0; i < n; ++i) ims.push_back(loadImage(files[i]));
  const std::string refName = p.getStringExpandTilde("reference");
  const Image refIm = loadImage(refName);
  const bool refNoCrop = p.checkAndGetBool("dontResizeOrCropReference");
  bool prepare = p.checkAndGetBool("cropInput") || p.checkAndGetBool("resizeInput") || refIm.nx != N || refIm.ny != N;
  for (const Image &im : ims) prepare = prepare || im.nx != N || im.ny != N;
  std::unique_ptr<DeviceBuffer> rawTpl;
  if (!prepare) {
    for (int i = 0; i < n; ++i) upload((char *)tpl.p + (size_t)i * es * nn, ims[i].data.data(), nn, dt);
    upload(ref.p, refIm.data.data(), nn, dt);
  } else {
    const bool scale = !p.checkAndGetBool("dontNormalizeInputImages");
    rawTpl.reset(new DeviceBuffer(es * nn * n));
    for (int i = 0; i < n; ++i) {
      prepareFrame(p, ims[i], files[i], L, dt, scale, true, false, (char *)tpl.p + (size_t)i * es * nn);
      prepareFrame(p, ims[i], files[i], L, dt, false, false, false, (char *)rawTpl->p + (size_t)i * es * nn);
    }
    prepareFrame(p, refIm, refName, L, dt, scale, true, refNoCrop, ref.p);
    sp.dont_normalize = 1;
    sp.pre_smooth_sigma = 0;
    sp.raw_templates = rawTpl->p;
  }
  ims.clear();
  std::unique_ptr<DeviceBuffer> pin;
  if (loadStageOneResults && n > 1) {
    // pairwise deformation of template i is stage 1's deformation of its template i-1
    pin.reset(new DeviceBuffer(es * 2 * nn * n));
    for (int i = 1; i < n; ++i) {
      std::vector<double> phi;
      // with altStartLevel the kept result, else the start-level solve's (matchSeries.h:206-211)
      if (useAltStartLevel(p))
        loadField(strprintf("%s../stage1/deformation-%03d_%%d.dat.bz2", dir.c_str(), i - 1), N, phi);
      else
        loadField(strprintf("%s../stage1/%d/deformation_%02d_%%d.dat.bz2", dir.c_str(), i - 1, L), N, phi);
      upload((char *)pin->p + (size_t)i * es * 2 * nn, phi.data(), 2 * nn, dt);
    }
  }
  std::vector<double> energies(n);
  // per-level records: every level of every solve, as the reference writes them
  const bool accumulate = !p.checkAndGetBool("dontAccumulateDeformation");
  const int npw = pin ? 1 : n;
  const int nlp = sp.stop_level - sp.start_level + 1, nlr = sp.refine_stop_level - sp.refine_start_level + 1;
  const bool inverse = accumulate && sp.calc_inverse;
  const int nla = sp.alt_start_level > 0 ? sp.stop_level - sp.alt_start_level + 1 : 0;
  const int nrec = npw * nlp + (accumulate ? (n - 1) * nlr : 0) + (inverse ? n * nlr : 0) + npw * nla;
  std::unique_ptr<DeviceBuffer> first;
  if (nla > 0) {
    first.reset(new DeviceBuffer(es * 2 * nn * n));
    sp.first_try_out = first->p;
  }
  // every level record's directory and level, in the library's record order (msg_series_levels)
  std::vector<LevelWriter::Slot> slots;
  slots.reserve(nrec);
  auto addSolve = [&](const std::string &sub, int l0, int l1) {
    makeDirectory(sub);
    for (int l = l0; l <= l1; ++l) slots.push_back({sub, l});
  };
  for (int i = 0; i < npw; ++i) addSolve(strprintf("%s%d/", dir.c_str(), i), sp.start_level, sp.stop_level);
  for (int i = 1; accumulate && i < n; ++i)
    addSolve(strprintf("%s%d-r/", dir.c_str(), i), sp.refine_start_level, sp.refine_stop_level);
  for (int i = 0; inverse && i < n; ++i)
    addSolve(strprintf("%s%d-r-inv/", dir.c_str(), i), sp.refine_start_level, sp.refine_stop_level);
  for (int i = 0; i < npw && nla > 0; ++i) addSolve(strprintf("%s%d-alt/", dir.c_str(), i), sp.alt_start_level, sp.stop_level);
  if ((int)slots.size() != nrec) MSH_THROW("internal error: level record count");
  LevelWriter writer(std::move(slots), sp, 8);
  msg_series_levels lv = {nullptr, nullptr, nullptr, nullptr, &LevelWriter::onLevel, &writer};
  const auto t0 = std::chrono::steady_clock::now();
  DeviceBuffer reduced(es * 2 * nn);
  msgCheck(msg_match_series(&sp, dt, tpl.p, n, ref.p, pin ? pin->p : nullptr, defs.p, pw.p, energies.data(), mean.p,
                            cnt.p, med.p, &lv, sp.reduce_deformations ? reduced.p : nullptr));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cerr << "Registered " << n << " templates on the GPU in " << secs << " s\n";
  writer.finish();

  std::vector<double> hdefs(2 * nn * n), hpw(2 * nn * n), hfirst(first ? 2 * nn * n : 0);
  download(hdefs.data(), defs.p, 2 * nn * n, dt);
  download(hpw.data(), pw.p, 2 * nn * n, dt);
  if (first) download(hfirst.data(), first->p, 2 * nn * n, dt);
  std::ofstream defNorms, energiesFile(dir + "/energies.txt");
  if (!loadStageOneResults) defNorms.open(dir + "/defNorms.txt");
  const int refineStop = sp.refine_stop_level;
  for (int i = 0; i < n; ++i) {
    const double *phw = hpw.data() + (size_t)i * 2 * nn;
    if (!loadStageOneResults || i == 0) {
      // the start-level solve's result (with altStartLevel the kept one may be the repeated solve's)
      const double *phs = first ? hfirst.data() + (size_t)i * 2 * nn : phw;
      const std::string sub = strprintf("%s%d/", dir.c_str(), i);
      if (sp.stop_level < L) saveField(sub + strprintf("deformation_%02d_%%d.dat.bz2", L), phs, N);
      if (!loadStageOneResults) {
        double norm = 0;
        for (int c = 0; c < 2; ++c) {
          double s = 0;
          for (size_t k = 0; k < nn; ++k) s += phs[c * nn + k] * phs[c * nn + k];
          norm += std::sqrt(s);
        }
        // matchSeismicSeries: the registration object's own deformation is never set on its dyadic
        // path, so its norm is 0 (NonDyadicRegistrationMultilevelDescent::getTransformationNorm)
        defNorms << i << " " << (kSeismic ? 0.0 : norm) << " # " << files[i] << "\n";
      }
    }
    if (accumulate && i > 0) {
      const std::string sub = strprintf("%s%d-r/", dir.c_str(), i);
      const double *phd = hdefs.data() + (size_t)i * 2 * nn;
      if (refineStop < L) saveField(sub + strprintf("deformation_%02d_%%d.dat.bz2", L), phd, N);
    }
    // matchSeismicSeries: getEnergyOfLastSolution is set by the full-resolution step only (NaN here)
    energiesFile << i << " " << (kSeismic ? (double)NAN : energies[i]) << "\n";
  }
  for (int i = 0; nla > 0 && i < npw; ++i)  // the kept results of the repeated solves
    saveField(strprintf("%sdeformation-%03d_%%d.dat.bz2", dir.c_str(), i), hpw.data() + (size_t)i * 2 * nn, N);
  if (sp.reduce_deformations) {  // "<saveDirectory>reduceDef_<c>.dat.bz2" (matchSeries.h:701)
    std::vector<double> hr(2 * nn);
    download(hr.data(), reduced.p, 2 * nn, dt);
    saveField(dir + "reduceDef_%d.dat.bz2", hr.data(), N);
  }
  std::vector<double> hm(nn), hc(nn), hmed(nn);
  download(hm.data(), mean.p, nn, dt);
  download(hc.data(), cnt.p, nn, dt);
  download(hmed.data(), med.p, nn, dt);
  writeReconstruction(dir, hm, hmed, hc, N);
  seriesExtras(p, stage, files, rawTpl ? rawTpl->p : tpl.p, defs.p, sp.reduce_deformations ? reduced.p : nullptr, N,
               dt);
}

// aol::MixedFormat(3, 6), the format aol::Vec prints its entries with (aol.cpp:167-216): fixed
// "%10.6Lf", or "%.2e" after a blank (two for non-negative values) when the value rounded to 1e-6
// has an integer part of 100 or more, or none and a fraction below 1e-3.
std::string mixedFormat(double v) {
  const int before = 3, after = 6;
  const long double val = v;
  const long double mx = powl(10.0L, (long double)(before - 1)), eps = powl(10.0L, (long double)(-after));
  const long double rval = eps * rintl(val / eps);
  long double l;
  const long double r = fabsl(modfl(rval, &l));
  std::ostringstream oss;
  if (std::isinf(v)) {
    oss << std::string(before - 2, ' ') << std::setw(2 + 1 + after) << std::left << (v > 0 ? " INF" : "-INF");
  } else if (std::isnan(v)) {
    oss << std::string(before - 1, ' ') << std::setw(1 + 1 + after) << std::left << "NAN";
  } else if (v == 0.0) {
    oss << std::setw(before) << std::right << 0 << std::string(1 + after, ' ');
  } else if (fabsl(l) >= mx || (l == 0.0L && r < sqrtl(eps))) {
    oss << std::setprecision(after - 4) << std::left << std::scientific << std::string(before - 2, ' ')
        << (val >= 0.0L ? " " : "") << val;
  } else {
    oss << strprintf("%*.*Lf", before + after + 1, after, val);
  }
  return oss.str();
}

// aol::Vector::getSaturatedMinMaxValue (aol/vec.h:864-870) as setOverflowHandlingToEnhanceContrast
// uses it (aol/vec.h:342-345): the sorted values at int(p/200 (n-1)) and int((1 - p/200)(n-1)),
// the upper one raised to min + 1 unless above the lower.
void enhancedRange(const std::vector<double> &v, double pct, double &mn, double &mx) {
  std::vector<double> t(v);
  std::sort(t.begin(), t.end());
  const int m1 = (int)t.size() - 1;
  mn = t[std::min(std::max((int)(pct / 200 * m1), 0), m1)];
  const double hi = t[std::min(std::max((int)((1 - pct / 200) * m1), 0), m1)];
  mx = hi > mn ? hi : mn + 1;
}

// ACTION 2: SeriesMatching::analyzeDeformations (matchSeries.h:578-676) on the stage's saved
// deformations: per frame the mean displacement (the translation) and the largest and mean pointwise
// norm of the rest (the non-rigid component), with "saveDeformedTemplates" each template deformed by
// its deformation (defTempl_<iii>.dat.bz2, +Inf outside, and a PNG preview with the template's mean
// outside) and with "saveXDerivativeOfDef" the pointwise norm of the x derivative of the non-rigid
// component (def_<iii>.q2bz). The translation plots go through gnuplot as in the reference.
void analyzeDeformations(const ParameterParser &p) {
  const std::string dir = dirOf(p);
  const int L = p.getInt("precisionLevel"), N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = fieldType(p), es = dt == MSG_F64 ? 8 : 4;
  const double H = 1.0 / (N - 1);
  const std::vector<std::string> files = createTemplateFileNameList(p);
  const int n = (int)files.size();
  const bool saveDef = p.checkAndGetBool("saveDeformedTemplates");
  const int nearest = p.checkAndGetBool("saveDeformedTemplatesUsingNearestNeighborInterpolation") ? 1 : 0;
  std::vector<std::array<double, 2>> tr(n);
  double avgMax = 0, avgMean = 0;
  std::vector<double> phi, temp(nn), norms(nn);
  for (int i = 0; i < n; ++i) {
    loadField(deformationBase(p, dir, i, L) + "_%d.dat.bz2", N, phi);
    if (saveDef) {
      // loadAndPrepareImage(file, curTemplate, NoScaling = true): unnormalised, pre-smoothed
      DeviceBuffer tpl((size_t)es * nn), d((size_t)es * 2 * nn), out((size_t)es * nn);
      prepareFrame(p, loadImage(files[i]), files[i], L, dt, false, true, false, tpl.p);
      upload(d.p, phi.data(), 2 * nn, dt);
      msgCheck(msg_warp(dt, tpl.p, d.p, out.p, N, 1, 0, 0, INFINITY, nearest, nullptr));
      download(temp.data(), out.p, nn, dt);
      saveP9(strprintf("%sdefTempl_%03d.dat.bz2", dir.c_str(), i), temp.data(), N, N);
      std::vector<double> t(nn);
      download(t.data(), tpl.p, nn, dt);
      const double m = meanOf(t);
      for (size_t k = 0; k < nn; ++k)
        if (std::isinf(temp[k])) temp[k] = m;
      double mn, mx;
      if (p.hasVariable("enhanceContrastSaturationPercentage")) {
        enhancedRange(temp, p.getDouble("enhanceContrastSaturationPercentage"), mn, mx);
      } else {
        mn = *std::min_element(temp.begin(), temp.end());
        mx = *std::max_element(temp.begin(), temp.end());
      }
      savePNGClipScaled(strprintf("%sdefTempl_%03d-%f_%f.png", dir.c_str(), i, mn, mx), temp.data(), N, N, mn, mx);
    }
    for (int c = 0; c < 2; ++c) {  // qc::MultiArray::getMeanValue: sum in order / size
      double s = 0;
      for (size_t k = 0; k < nn; ++k) s += phi[c * nn + k];
      tr[i][c] = s / (double)nn;
      for (size_t k = 0; k < nn; ++k) phi[c * nn + k] += -tr[i][c];
    }
    for (size_t k = 0; k < nn; ++k) {
      double s2 = 0;
      s2 += phi[k] * phi[k];
      s2 += phi[nn + k] * phi[nn + k];
      norms[k] = std::sqrt(s2);
    }
    const double nmax = *std::max_element(norms.begin(), norms.end());
    double s = 0;
    for (size_t k = 0; k < nn; ++k) s += norms[k];
    avgMax += nmax;
    avgMean += s / (double)nn;
    double t2 = 0;
    t2 += tr[i][0] * tr[i][0];
    t2 += tr[i][1] * tr[i][1];
    std::cerr << "Translation " << mixedFormat(tr[i][0]) << " " << mixedFormat(tr[i][1])
              << ", non-rigid component max " << nmax / std::sqrt(t2) * 100 << "%, " << nmax / H << " pixels"
              << std::endl;
    if (p.checkAndGetBool("saveXDerivativeOfDef")) {  // ScalarArrayHelper::getPointWiseXDerivNorm
      for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
          double v2 = 0;
          for (int c = 0; c < 2; ++c) {
            const double *a = phi.data() + c * nn + (size_t)y * N;
            const double dx = x == 0 ? a[1] - a[0] : (x == N - 1 ? a[x] - a[x - 1] : 0.5 * (a[x + 1] - a[x - 1]));
            const double g = dx / H;
            v2 += g * g;
          }
          temp[(size_t)y * N + x] = std::sqrt(v2);
        }
      saveP9(strprintf("%sdef_%03d.q2bz", dir.c_str(), i), temp.data(), N, N);
    }
  }
  avgMax /= n;
  avgMean /= n;
  double avgTr = 0;
  for (int i = 1; i < n; ++i) {
    const double a = tr[i][0] - tr[i - 1][0], b = tr[i][1] - tr[i - 1][1];
    double q = 0;
    q += a * a;
    q += b * b;
    avgTr += std::sqrt(q);
  }
  avgTr /= (n - 1);
  std::cerr << "average translation = " << avgTr << ", " << avgTr / H << " pixels" << std::endl;
  std::cerr << "average non-rigid component = " << avgMean << ", " << avgMean / H << " pixels" << std::endl;
  std::cerr << "ratio  = " << avgMean / avgTr << std::endl;
  std::cerr << "average non-rigid component max = " << avgMax << ", " << avgMax / H << " pixels" << std::endl;
  std::cerr << "ratio  = " << avgMax / avgTr << std::endl;
  // <dir>trans.png (the translations as a curve) and <dir>norms.png (their norms over the frame
  // number), drawn by gnuplot as aol::Plotter does (aol/gnuplotter.h)
  auto plot = [&](const std::string &base, bool curve) {
    const std::string dat = base + ".dat", gp = base + ".gnuplot";
    {
      std::ofstream f(dat);
      f.precision(17);
      for (int i = 0; i < n; ++i) {
        if (curve) f << tr[i][0] << " " << tr[i][1] << "\n";
        else f << (i + 1) << " " << std::sqrt(tr[i][0] * tr[i][0] + tr[i][1] * tr[i][1]) << "\n";
      }
    }
    {
      std::ofstream f(gp);
      f << "set terminal png\nset output \"" << base << ".png\"\nplot \"" << dat << "\" with lines notitle\n";
    }
    if (std::system(("gnuplot \"" + gp + "\"").c_str()) != 0)
      std::cerr << "aol::runGnuplot: Calling gnuplot returned an error." << std::endl;
    std::remove(dat.c_str());
    std::remove(gp.c_str());
  };
  plot(dir + "trans", true);
  plot(dir + "norms", false);
}

// ACTION 3: SeriesMatching::applyDeformation (matchSeries.h:706-734): the sum of the saved
// deformations <deformationBaseNamePattern with i, i> for i = numDeformations-1 down to 0
// (setTransformation + addTransformationTo: a plain sum), applied to every template (0 or the
// template's mean outside, optionally nearest neighbour) and written as <basename>.q2bz (or into a
// copy of the DM3/DM4 container). Stage 1 puts the reference in front of the templates; the loop
// keeps the count of templates without it, as the reference does.
void applyDeformation(const ParameterParser &p, int stage) {
  const std::string dir = dirOf(p);
  const int L = p.getInt("precisionLevel"), N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = fieldType(p), es = dt == MSG_F64 ? 8 : 4;
  std::vector<std::string> files = createTemplateFileNameList(p);
  const int count = (int)files.size();
  if (stage == 1) files.insert(files.begin(), p.getStringExpandTilde("reference"));
  const int numDef = p.getInt("numDeformations");
  const std::string pat = p.getString("deformationBaseNamePattern");
  std::vector<double> sum, phi;
  loadField(strprintf(pat.c_str(), numDef - 1, numDef - 1) + "_%d.dat.bz2", N, sum);
  for (int i = numDef - 2; i >= 0; --i) {
    loadField(strprintf(pat.c_str(), i, i) + "_%d.dat.bz2", N, phi);
    for (size_t k = 0; k < 2 * nn; ++k) sum[k] += phi[k];
  }
  DeviceBuffer d((size_t)es * 2 * nn), tpl((size_t)es * nn), out((size_t)es * nn);
  upload(d.p, sum.data(), 2 * nn, dt);
  const bool extMean = p.checkAndGetBool("saveNamedDeformedTemplatesExtendedWithMean");
  const int nearest = p.checkAndGetBool("saveNamedDeformedTemplatesUsingNearestNeighborInterpolation") ? 1 : 0;
  std::vector<double> h(nn);
  for (int i = 0; i < count; ++i) {
    // loadAndPrepareImage(file, curTemplate, NoScaling, NoResizeOrCrop = false, NoSmoothing)
    loadRaw(p, files[i], L, dt, tpl.p);
    double ext = 0;
    if (extMean) {
      download(h.data(), tpl.p, nn, dt);
      ext = meanOf(h);
    }
    msgCheck(msg_warp(dt, tpl.p, d.p, out.p, N, 1, 0, 0, ext, nearest, nullptr));
    download(h.data(), out.p, nn, dt);
    saveNamedDeformedTemplate(p, files[i], h.data(), N);
  }
}

void doAction(ParameterParser &p, int action, bool loadStageOneResults, int stage) {
  // the registration object's constructor prepares the stage directory (registration.h:263-274)
  makeDirectory(dirOf(p));
  p.dumpToFile("parameter-dump.txt", dirOf(p));
  switch (action) {
    case 0: matchAndAverage(p, loadStageOneResults, stage); break;
    case 1: averageOnly(p, stage); break;
    case 2: analyzeDeformations(p); break;
    case 3: applyDeformation(p, stage); break;
    default: MSH_THROW("SeriesMatching::doAction: Invalid ActionType");
  }
}

// matchSeries<ConfType, RegisType> (matchSeries.h:1249-1320)
void runMatchSeries(ParameterParser &parser, int argc, char **argv) {
  const auto t0 = std::chrono::steady_clock::now();
  const int numExtraStages = parser.getInt("numExtraStages");
  const double extraStagesLambdaFactor = parser.getDouble("extraStagesLambdaFactor");
  const std::string saveDir = dirOf(parser);
  makeDirectory(saveDir);
  parser.dumpToFile("parameter-dump.txt", saveDir);
  std::ofstream logFile(saveDir + "log.txt");
  TeeBuf tee(std::cerr.rdbuf(), logFile.rdbuf());
  std::streambuf *old = std::cerr.rdbuf(&tee);
  struct Restore {
    std::streambuf *o;
    ~Restore() { std::cerr.rdbuf(o); }
  } restore{old};
  std::cerr << "matchSeries on " << msg_version() << "\n";

  if (parser.hasVariable("templateNamePattern")) {
    if (parser.hasVariable("reference"))
      MSH_THROW("Paramater \"reference\" may not be specified if \"templateNamePattern\" is used.");
    if (!parser.hasVariable("initialAverage")) {
      parser.addVariable("reference", strprintf(parser.getString("templateNamePattern").c_str(),
                                                parser.getInt("templateNumOffset")));
      parser.changeVariableValue("templateNumOffset", parser.getInt("templateNumOffset") + parser.getInt("templateNumStep"));
      parser.changeVariableValue("numTemplates", parser.getInt("numTemplates") - 1);
    } else {
      parser.addVariable("reference", parser.getStringExpandTilde("initialAverage"));
    }
  }
  if (!parser.hasVariable("template")) parser.addVariable("template", parser.getString("reference"));
  rejectUnsupported(parser);

  ParameterParser stage1(parser);
  stage1.changeVariableValue("saveDirectory", saveDir + "stage1/");
  const int action = argc > 2 ? atoi(argv[2]) : 0;
  if (!parser.checkAndGetBool("skipStage1")) doAction(stage1, action, false, 1);
  const std::string averageName = std::string(parser.checkAndGetBool("useMedianAsNewTarget") ? "median" : "average") + ".q2bz";
  std::string lastAverage = saveDir + "stage1/" + averageName;

  if (numExtraStages > 0) {
    ParameterParser stage2(parser);
    if (!parser.hasVariable("initialAverage")) {
      stage2.changeVariableValue("templateNumOffset", parser.getInt("templateNumOffset") - parser.getInt("templateNumStep"));
      stage2.changeVariableValue("numTemplates", parser.getInt("numTemplates") + 1);
    }
    // the reference passes lambda through "%f"
    stage2.changeVariableValue("lambda", strprintf("%f", parser.getDouble("lambda") * extraStagesLambdaFactor));
    if (parser.checkAndGetBool("cropInput") || parser.checkAndGetBool("resizeInput"))
      stage2.changeVariableValue("dontResizeOrCropReference", 1);
    if (parser.checkAndGetBool("useCorrelationToInitTranslation"))
      stage2.changeVariableValue("useCorrelationToInitTranslation", 0);
    for (int i = 0; i < numExtraStages; ++i) {
      ParameterParser stageN(stage2);
      stageN.changeVariableValue("saveDirectory", strprintf("%sstage%d/", saveDir.c_str(), i + 2));
      stageN.changeVariableValue("reference", lastAverage);
      stageN.addVariable("stage", i + 2);
      doAction(stageN, action, parser.checkAndGetBool("reuseStage1Results"), i + 2);
      lastAverage = strprintf("%sstage%d/", saveDir.c_str(), i + 2) + averageName;
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cerr << "--------------------------------------------------------------\n"
            << "Wall clock time = " << secs << " s\n"
            << "All output has been written to file " << saveDir << "log.txt\n";
}

}  // namespace

// matchSeismicSeries' front end (matchSeismicSeries.cpp:24-82): the save directory optionally named
// after the parameters, the initial average made by interleaving the reference's rows with the
// templates' (without alignment: "justStackIniAverage" or "justCopyLastRefLineForIniAverage"), and
// the dyadic path only: the full-resolution step of NonDyadicRegistrationMultilevelDescent and the
// aligned interleaving need a Cholesky factorisation of the H1 metric, which is not part of this build.
void seismicFrontEnd(ParameterParser &parser) {
  if (parser.checkAndGetBool("addParamsToSafeDirName")) {
    std::string d = parser.getStringExpandTilde("saveDirectory");
    if (!d.empty() && d.back() == '/') d.pop_back();
    std::ostringstream o;
    o << d;
    if (parser.checkAndGetBool("saveNamedDeformedTemplatesUsingNearestNeighborInterpolation")) o << "-nointerp";
    o << "-" << parser.getDouble("lambda") << "-" << parser.getDouble("yRegFactor") << "-"
      << parser.getDouble("laplaceWeight") << "-" << parser.getDouble("averageXDerivNormThreshold") << "-"
      << parser.getDouble("extraStagesLambdaFactor");
    if (parser.checkAndGetBool("forceNonDyadicXShiftYConst")) o << "-yconst";
    o << "/";
    parser.changeVariableValue("saveDirectory", o.str());
  }
  if (parser.checkAndGetBool("forceNonDyadicXShiftYConst"))
    MSH_THROW("\"forceNonDyadicXShiftYConst 1\" selects the full-resolution step, which this build does not run");
  if (parser.getDouble("laplaceWeight") > 0) MSH_THROW("\"laplaceWeight\" > 0 is not supported by this build");
  (void)parser.getInt("keepEmptyTopBottomBlocks");  // read by the registration object's constructor
  const int L = parser.getInt("precisionLevel");
  if (!parser.hasVariable("dyadicDepth") || parser.getInt("dyadicDepth") != L || parser.getInt("stopLevel") != L)
    MSH_THROW("matchSeismicSeries: this build runs the dyadic path only: \"dyadicDepth\", \"stopLevel\" and "
              "\"precisionLevel\" equal and the frames of 2^precisionLevel x 2^precisionLevel (the full-resolution "
              "step needs a Cholesky factorisation of the H1 metric)");
  const std::vector<std::string> files = createTemplateFileNameList(parser);
  for (const std::string &f : files) {
    const Image im = loadImage(f);
    if (im.nx != (1 << L) || im.ny != (1 << L))
      MSH_THROW(strprintf("matchSeismicSeries: %s is %dx%d; this build runs the dyadic path only (frames of "
                          "2^precisionLevel x 2^precisionLevel)", f.c_str(), im.nx, im.ny));
    for (int y = 0; y < im.ny; ++y) {
      double s2 = 0;
      for (int x = 0; x < im.nx; ++x) s2 += im.data[(size_t)y * im.nx + x] * im.data[(size_t)y * im.nx + x];
      if (s2 == 0) MSH_THROW(strprintf("matchSeismicSeries: %s has an empty row (%d); rows are dropped from the "
                                       "dyadic grid only by the full-resolution path", f.c_str(), y));
    }
  }
  if (!parser.hasVariable("initialAverage")) {
    // interleaveImages (matchSeismicSeries.h:9-96) without alignment
    const bool align = parser.getInt("justStackIniAverage") == 0;
    const bool copyLast = parser.hasVariable("justCopyLastRefLineForIniAverage") &&
                          parser.checkAndGetBool("justCopyLastRefLineForIniAverage");
    if (align && !copyLast)
      MSH_THROW("matchSeismicSeries: the aligned interleaving needs a Cholesky factorisation of the 1-D H1 metric; "
                "set \"justStackIniAverage 1\" or give \"initialAverage\"");
    const int numTemplates = parser.getInt("refLineDistance") - 1;
    if (numTemplates >= (int)files.size())
      MSH_THROW("refLineDistance too big for the amount of supplied template images");
    const bool clamp = parser.hasVariable("clampRange");
    auto load = [&](const std::string &f) {
      Image im = loadImage(f);
      if (clamp)
        for (double &v : im.data) v = std::min(std::max(v, parser.getDouble("clampRange", 0)), parser.getDouble("clampRange", 1));
      return im;
    };
    const Image ref = load(files[0]);
    std::vector<Image> tpls;
    for (int i = 0; i < numTemplates; ++i) tpls.push_back(load(files[i + 1]));
    std::vector<double> out(ref.data);
    const int nx = ref.nx, ny = ref.ny, dist = numTemplates + 1;
    auto rowEmpty = [&](const Image &im, int r) {
      double s2 = 0;
      for (int x = 0; x < nx; ++x) s2 += im.data[(size_t)r * nx + x] * im.data[(size_t)r * nx + x];
      return s2 == 0;
    };
    int lastNonEmpty = -1;
    for (int line = 0; line < ny; ++line) {
      if (rowEmpty(ref, line)) {
        if (line == 0 && numTemplates > 0)
          std::copy(tpls[0].data.begin(), tpls[0].data.begin() + nx, out.begin());
        else if (copyLast && lastNonEmpty >= 0)
          std::copy(ref.data.begin() + (size_t)lastNonEmpty * nx, ref.data.begin() + (size_t)(lastNonEmpty + 1) * nx,
                    out.begin() + (size_t)line * nx);
        continue;
      }
      lastNonEmpty = line;
      if (copyLast || line + dist >= ny) continue;
      if (rowEmpty(ref, line + dist))
        MSH_THROW(strprintf("Unexpected empty line in reference (current line index is %d)", line));
      for (int i = 0; i < numTemplates; ++i) {
        if (rowEmpty(tpls[i], line + i + 1))
          MSH_THROW(strprintf("Unexpected empty line in template (current line index is %d)", line));
        std::copy(tpls[i].data.begin() + (size_t)(line + i + 1) * nx, tpls[i].data.begin() + (size_t)(line + i + 2) * nx,
                  out.begin() + (size_t)(line + 1 + i) * nx);
      }
    }
    std::string d = parser.getStringExpandTilde("saveDirectory");
    makeDirectory(d);
    const std::string name = d + "initialAverage.q2bz";
    saveP9(name, out.data(), nx, ny);
    parser.addVariable("initialAverage", name);
  }
}

int main(int argc, char **argv) {
  const char *strictEnv = getenv("MSG_STRICT");
  const bool strict = strictEnv && atoi(strictEnv) == 1;
  try {
    ParameterParser parser(argc > 1 ? argv[1] : "matchSeries.par");
    if (kSeismic) {
      seismicFrontEnd(parser);
    } else {
      const int model = parser.getIntOrDefault("deformationModel", 0);
      if (model != 22 && model != 0)
        MSH_THROW(strprintf("deformationModel %d is not supported by this build (0: NCC, 22: SSD; both with the "
                            "Dirichlet regulariser)", model));
    }
    runMatchSeries(parser, argc, argv);
  } catch (Exception &e) {
    e.dump(std::cerr);
    return strict ? 1 : 0;
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return strict ? 1 : 0;
  }
  return 0;
}
