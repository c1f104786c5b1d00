// Golden-vector harness around the reference implementation (test infrastructure only).
//
// This file is compiled by oracle/Makefile against the reference headers and sources where they
// lie (QuocMesh modules aol/, quoc/, imgproc/); the binary lands in oracle/_ref/ and is only run by
// tests/golden/make_golden.py in the build container to write small fixtures. Nothing in the
// product links or calls it.
//
// Protocol: ref_harness <cmd> <N> [params...] <in.bin> <out.bin>
//   in.bin / out.bin are raw little-endian float64 arrays, concatenated in the order listed per
//   command below. N is the number of nodes per side (2^level).

#include <registration.h>
#include <cellCenteredGrid.h>
#include <deformations.h>
#include <hyperelastic.h>
#include <linearSmoothOp.h>
#include <gradientDescent.h>
#include <multilevelArray.h>
#include <AmbrosioTortorelli.h>
#include <matchSeries.h>

#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <vector>

typedef double RType;
typedef qc::CellCenteredCubicGrid<qc::QC_2D> GridType;
typedef qc::RectangularGridConfigurator<RType, qc::QC_2D, aol::GaussQuadrature<RType, qc::QC_2D, 3>, GridType> ConfType;
typedef qc::LinearSmoothOp<RType, qc::CellCenteredGridTrait<RType, qc::QC_2D> > SmoothType;

static std::vector<double> readAll(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  fseek(f, 0, SEEK_END);
  long n = ftell(f) / 8;
  fseek(f, 0, SEEK_SET);
  std::vector<double> v(n);
  if (n && fread(v.data(), 8, n, f) != (size_t)n) { fprintf(stderr, "short read\n"); exit(2); }
  fclose(f);
  return v;
}

static void writeAll(const char *path, const std::vector<double> &v) {
  FILE *f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  fwrite(v.data(), 8, v.size(), f);
  fclose(f);
}

static int levelOf(int N) {
  int l = 0;
  while ((1 << l) < N) ++l;
  if ((1 << l) != N) { fprintf(stderr, "N must be a power of two\n"); exit(2); }
  return l;
}

static void toVec(const double *src, aol::Vector<RType> &dst) {
  for (int i = 0; i < dst.size(); ++i) dst[i] = src[i];
}
static void append(std::vector<double> &out, const aol::Vector<RType> &v) {
  for (int i = 0; i < v.size(); ++i) out.push_back(v[i]);
}

// Gradient descent type used by StandardRegistration::updateTransformationWithGD for model 22,
// subclassed only to record the step size chosen each iteration and the energy evaluations.
typedef aol::H1GradientDescent<ConfType, aol::MultiVector<RType>, SmoothType> H1Type;
typedef aol::GradientDescentWithAutomaticFilterWidth<ConfType, aol::MultiVector<RType>, H1Type> GDType;

class TracingGD : public GDType {
public:
  mutable std::vector<double> taus;      // every Armijo result, in call order
  mutable std::vector<double> sigmas;    // filter width at each Armijo call
  mutable int evals;
  mutable std::vector<int> evalsPerIter;
  TracingGD(const GridType &g, const aol::Op<aol::MultiVector<RType>, aol::Scalar<RType> > &E,
            const aol::Op<aol::MultiVector<RType> > &DE, int maxIt, RType tau, RType eps)
      : GDType(g, E, DE, maxIt, tau, eps), evals(0) {}
protected:
  RType getStepSizeWithSelectedController(aol::MultiVector<RType> &Dir, const aol::MultiVector<RType> &Pos,
                                          const RType OldTau) const {
    const int before = evals;
    RType t = GDType::getStepSizeWithSelectedController(Dir, Pos, OldTau);
    taus.push_back(t);
    sigmas.push_back(this->_filterWidth);
    evalsPerIter.push_back(evals - before);
    return t;
  }
  RType ArmijoLineSearchHelpFunction_evaluate(const aol::MultiVector<RType> &Dir, const aol::MultiVector<RType> &Pos,
                                              const RType tw) const {
    if (tw != 0) ++evals;
    return GDType::ArmijoLineSearchHelpFunction_evaluate(Dir, Pos, tw);
  }
};

// The seismic x-shift solve of findXShift (matchSeries.h:805-868) on the dyadic grid: a scalar field
// with the V-cycle metric and the automatic filter width, traced like TracingGD.
typedef qc::RectangularGridConfigurator<RType, qc::QC_1D, aol::GaussQuadrature<RType, qc::QC_1D, 3> > Conf1D;
typedef qc::FELineShiftRegisEnergy<ConfType, Conf1D, qc::NCCRegistrationConfigurator<Conf1D> > LineEnergy;
typedef aol::H1GradientDescent<ConfType, aol::Vector<RType>, SmoothType> H1VType;
typedef aol::GradientDescentWithAutomaticFilterWidth<ConfType, aol::Vector<RType>, H1VType> GDVType;

class TracingGDV : public GDVType {
public:
  mutable std::vector<double> taus, sigmas;
  mutable int evals;
  mutable std::vector<int> evalsPerIter;
  TracingGDV(const GridType &g, const aol::Op<aol::Vector<RType>, aol::Scalar<RType> > &E,
             const aol::Op<aol::Vector<RType> > &DE, int maxIt)
      : GDVType(g, E, DE, maxIt), evals(0) {}
protected:
  RType getStepSizeWithSelectedController(aol::Vector<RType> &Dir, const aol::Vector<RType> &Pos,
                                          const RType OldTau) const {
    const int before = evals;
    RType t = GDVType::getStepSizeWithSelectedController(Dir, Pos, OldTau);
    taus.push_back(t);
    sigmas.push_back(this->_filterWidth);
    evalsPerIter.push_back(evals - before);
    return t;
  }
  RType ArmijoLineSearchHelpFunction_evaluate(const aol::Vector<RType> &Dir, const aol::Vector<RType> &Pos,
                                              const RType tw) const {
    if (tw != 0) ++evals;
    return GDVType::ArmijoLineSearchHelpFunction_evaluate(Dir, Pos, tw);
  }
};

int main(int argc, char **argv) {
  if (argc < 5) { fprintf(stderr, "usage: ref_harness cmd N [params] in out\n"); return 2; }
  try {
    const std::string cmd = argv[1];
    const int N = atoi(argv[2]);
    const int L = levelOf(N);
    const int NN = N * N;
    const char *inPath = argv[argc - 2];
    const char *outPath = argv[argc - 1];
    std::vector<double> in = readAll(inPath);
    std::vector<double> out;
    GridType grid(L, qc::QC_2D);

    if (cmd == "ssd") {
      // in: R, T, phix, phiy ; out: E_ssd, E_reg, DE_ssd (2N^2), K phi (2N^2)
      aol::Vector<RType> R(NN), T(NN);
      aol::MultiVector<RType> phi(2, NN), dssd(2, NN), kphi(2, NN);
      toVec(&in[0], R); toVec(&in[NN], T); toVec(&in[2 * NN], phi[0]); toVec(&in[3 * NN], phi[1]);
      qc::SSDEnergy<ConfType> E(grid, R, T);
      qc::SSDForce<ConfType> F(grid, R, T);
      aol::StiffOp<ConfType> stiff(grid, aol::ASSEMBLED);
      qc::DisplacementLengthEnergy<ConfType> regE(stiff);
      aol::DiagonalBlockOp<RType> regDE(stiff);
      aol::Scalar<RType> e;
      E.apply(phi, e);
      out.push_back(e[0]);
      regE.apply(phi, e);
      out.push_back(e[0]);
      F.apply(phi, dssd);
      regDE.apply(phi, kphi);
      append(out, dssd[0]); append(out, dssd[1]); append(out, kphi[0]); append(out, kphi[1]);
    } else if (cmd == "ncc") {
      // in: R, T, phix, phiy ; out: E_ncc, DE_ncc (2N^2), M R (N^2), normalised R (N^2), var(R)
      // (NormalizedCrossCorrelationEnergy / VariationOf... registration.h:1909-2018; MassOp assembled)
      aol::Vector<RType> R(NN), T(NN), mr(NN), rn(NN);
      aol::MultiVector<RType> phi(2, NN), dncc(2, NN);
      toVec(&in[0], R); toVec(&in[NN], T); toVec(&in[2 * NN], phi[0]); toVec(&in[3 * NN], phi[1]);
      qc::NormalizedCrossCorrelationEnergy<ConfType> E(grid, R, T);
      qc::VariationOfNormalizedCrossCorrelationEnergy<ConfType> DE(grid, R, T);
      aol::Scalar<RType> e;
      E.apply(phi, e);
      out.push_back(e[0]);
      DE.apply(phi, dncc);
      append(out, dncc[0]); append(out, dncc[1]);
      E.getMassOpRef().apply(R, mr);
      append(out, mr);
      const RType var = qc::NormalizedCrossCorrelationEnergy<ConfType>::normalizeImageForNCC(E.getMassOpRef(), R, rn);
      append(out, rn);
      out.push_back(var);
    } else if (cmd == "consistency") {
      // in: psi (2N^2), phi (2N^2) ; out: E, DE (2N^2) of ConsistencyEnergyOp / VariationOf...
      aol::MultiVector<RType> psi(2, NN), phi(2, NN), de(2, NN);
      toVec(&in[0], psi[0]); toVec(&in[NN], psi[1]); toVec(&in[2 * NN], phi[0]); toVec(&in[3 * NN], phi[1]);
      qc::ConsistencyEnergyOp<ConfType> E(grid, psi);
      qc::VariationOfConsistencyEnergyOp<ConfType> DE(grid, psi);
      aol::Scalar<RType> e;
      E.apply(phi, e);
      out.push_back(e[0]);
      DE.apply(phi, de);
      append(out, de[0]); append(out, de[1]);
    } else if (cmd == "reduce") {
      // params: n maxIt stopEps ; in: n deformations (2N^2 each) ; out: phi (2N^2), E, iterations
      // (SeriesMatching::reduceDeformations, matchSeries.h:678-704)
      const int n = atoi(argv[3]);
      const int maxIt = atoi(argv[4]);
      const double stopEps = atof(argv[5]);
      std::vector<aol::MultiVector<RType> *> defs;
      aol::RandomAccessContainer<const qc::ConsistencyEnergyOp<ConfType> > energies;
      aol::RandomAccessContainer<const qc::VariationOfConsistencyEnergyOp<ConfType> > derivatives;
      aol::LinCombOp<aol::MultiVector<RType>, aol::Scalar<RType> > E;
      aol::LinCombOp<aol::MultiVector<RType> > DE;
      for (int i = 0; i < n; ++i) {
        defs.push_back(new aol::MultiVector<RType>(2, NN));
        toVec(&in[(2 * i) * NN], (*defs[i])[0]);
        toVec(&in[(2 * i + 1) * NN], (*defs[i])[1]);
      }
      for (int i = 0; i < n; ++i) {
        energies.constructDatumAndPushBack(grid, *defs[i]);
        E.appendReference(energies[i]);
        derivatives.constructDatumAndPushBack(grid, *defs[i]);
        DE.appendReference(derivatives[i]);
      }
      typedef aol::H1GradientDescent<ConfType, aol::MultiVector<RType>, SmoothType> PlainH1;
      PlainH1 solver(grid, E, DE, maxIt, 1, stopEps);
      solver.setConfigurationFlags(PlainH1::USE_NONLINEAR_CG | PlainH1::LOG_GRADIENT_NORM_AT_OLD_POSITION |
                                   PlainH1::USE_GRADIENT_BASED_STOPPING | PlainH1::DO_NOT_WRITE_CONSOLE_OUTPUT);
      aol::MultiVector<RType> phi(2, NN);
      solver.applySingle(phi);
      append(out, phi[0]); append(out, phi[1]);
      out.push_back(solver.getEnergyAtLastPosition());
      out.push_back(solver.getIterations());
      for (auto *d : defs) delete d;
    } else if (cmd == "smooth") {
      // params: sigma ; in: d ; out: one V-cycle of the heat operator applied to d
      const double sigma = atof(argv[3]);
      aol::Vector<RType> d(NN), r(NN);
      toVec(&in[0], d);
      SmoothType s(grid);
      s.setSigma(sigma);
      s.apply(d, r);
      append(out, r);
    } else if (cmd == "warp") {
      // params: extconst nearest ; in: img, phix, phiy ; out: DeformImage(extend), FastDeformImage(clamp)
      const double ext = atof(argv[3]);
      const bool nearest = atoi(argv[4]) != 0;
      aol::Vector<RType> img(NN), a(NN), b(NN);
      aol::MultiVector<RType> phi(2, NN);
      toVec(&in[0], img); toVec(&in[NN], phi[0]); toVec(&in[2 * NN], phi[1]);
      qc::DeformImage<ConfType>(img, grid, a, phi, true, ext, nearest);
      qc::FastDeformImage<ConfType, qc::QC_2D>(img, grid, b, phi);
      append(out, a); append(out, b);
    } else if (cmd == "transfer") {
      // in: fine (N^2), coarse ((N/2)^2) ; out: restrict(fine), prolong(coarse), resample(coarse->N)
      const int n = N / 2;
      GridType cgrid(L - 1, qc::QC_2D);
      aol::Vector<RType> fine(NN), coarse(n * n), r(n * n), p(NN);
      toVec(&in[0], fine); toVec(&in[NN], coarse);
      qc::CellCenteredRestrictOp<RType, qc::QC_2D>(cgrid, grid).apply(fine, r);
      qc::CellCenteredProlongOp<RType, qc::QC_2D>(cgrid, grid).apply(coarse, p);
      qc::ScalarArray<RType, qc::QC_2D> src(n, n), dst(N, N);
      for (int i = 0; i < n * n; ++i) src[i] = coarse[i];
      dst.resampleFrom(src);
      append(out, r); append(out, p); append(out, dst);
    } else if (cmd == "compose") {
      // in: phi1 (2N^2), phi2 (2N^2) ; out: ConcatAndSmoothlyExtendDeformations(phi1, phi2)
      aol::MultiVector<RType> p1(2, NN), p2(2, NN), res(2, NN);
      toVec(&in[0], p1[0]); toVec(&in[NN], p1[1]); toVec(&in[2 * NN], p2[0]); toVec(&in[3 * NN], p2[1]);
      qc::ConcatAndSmoothlyExtendDeformations<ConfType>(p1, p2, grid, res);
      append(out, res[0]); append(out, res[1]);
    } else if (cmd == "resize") {
      // params: w h ; in: a w x h image ; out: its resampleFrom onto N x N
      const int w = atoi(argv[3]), h = atoi(argv[4]);
      qc::ScalarArray<RType, qc::QC_2D> src(w, h), dst(N, N);
      for (int i = 0; i < w * h; ++i) src[i] = in[i];
      dst.resampleFrom(src);
      append(out, dst);
    } else if (cmd == "inverse") {
      // in: phi (2N^2) ; out: qc::approxInverseDeformation of phi
      aol::MultiVector<RType> phi(2, NN), inv(2, NN);
      toVec(&in[0], phi[0]); toVec(&in[NN], phi[1]);
      qc::approxInverseDeformation<ConfType>(grid, phi, inv);
      append(out, inv[0]); append(out, inv[1]);
    } else if (cmd == "normalize") {
      // params: pct (negative: no saturation) normMinToZero ; in: img ; out: prepared img + pyramid
      const double pct = atof(argv[3]);
      const bool minToZero = atoi(argv[4]) != 0;
      qc::ScalarArray<RType, qc::QC_2D> img(N, N);
      for (int i = 0; i < NN; ++i) img[i] = in[i];
      if (pct >= 0) {
        const aol::Vec2<RType> mm = img.getSaturatedMinMaxValue(pct);
        img.clamp(mm[0], mm[1]);
        img.scaleValuesTo01();
      } else {
        RType minValue = img.getMinValue();
        if ((minValue < 0.) || minToZero) img.addToAll(-minValue);
        RType maxValue = img.getMaxValue();
        if (maxValue != 0.) img /= maxValue;
      }
      qc::MultilevelArray<RType, qc::ScalarArray<RType, qc::QC_2D>, qc::CellCenteredProlongOp<RType, qc::QC_2D>,
                          qc::CellCenteredRestrictOp<RType, qc::QC_2D>, GridType> mla(grid);
      mla.current() = img;
      mla.levRestrict(0, L);
      for (int l = L; l >= 0; --l) append(out, mla[l]);
    } else if (cmd == "median") {
      // params: F fill ; in: F stacked images (Inf = missing) ; out: mean, numSamples, median
      const int F = atoi(argv[3]);
      const double fill = atof(argv[4]);
      aol::MultiVector<RType> stack(F, NN);
      for (int f = 0; f < F; ++f) toVec(&in[f * NN], stack[f]);
      aol::Vector<RType> mean(NN), cnt(NN), med(NN);
      for (int j = 0; j < NN; ++j) {
        mean[j] = 0; int c = 0;
        for (int f = 0; f < F; ++f)
          if (stack[f][j] != aol::NumberTrait<RType>::Inf) { mean[j] += stack[f][j]; ++c; }
        if (c > 0) mean[j] /= c; else mean[j] = fill;
        cnt[j] = c;
      }
      stack.getMedianVecOverComponents(med, fill);
      append(out, mean); append(out, cnt); append(out, med);
    } else if (cmd == "load") {
      // params: file ; in: ignored ; out: nx, ny, then the array loaded by qc::ScalarArray<QC_2D>(file)
      qc::ScalarArray<RType, qc::QC_2D> a(argv[3]);
      out.push_back(a.getNumX());
      out.push_back(a.getNumY());
      append(out, a);
    } else if (cmd == "solve") {
      // params: lambda maxIt stopEps [deformationModel: 22 SSD (default), 0 NCC] ; in: R, T, phix0, phiy0
      // out: phix, phiy, iterations, E_final, then per iteration k: E_total(k), tau(k), sigma(k), evals(k)
      // (E_total(0) is the initial energy; tau/sigma/evals are recorded per Armijo call).
      const double lambda = atof(argv[3]);
      const int maxIt = atoi(argv[4]);
      const double stopEps = atof(argv[5]);
      aol::Vector<RType> R(NN), T(NN);
      aol::MultiVector<RType> phi(2, NN);
      toVec(&in[0], R); toVec(&in[NN], T); toVec(&in[2 * NN], phi[0]); toVec(&in[3 * NN], phi[1]);
      const int model = argc > 8 ? atoi(argv[6]) : 22;  // deformationModel: 22 SSD, 0 NCC
      aol::StiffOp<ConfType> stiff(grid, aol::ASSEMBLED);
      qc::DisplacementLengthEnergy<ConfType> regE(stiff);
      aol::DiagonalBlockOp<RType> regDE(stiff);
      qc::SSDEnergy<ConfType> ssdE(grid, R, T);
      qc::SSDForce<ConfType> ssdF(grid, R, T);
      qc::NormalizedCrossCorrelationEnergy<ConfType> nccE(grid, R, T);
      qc::VariationOfNormalizedCrossCorrelationEnergy<ConfType> nccDE(grid, R, T);
      aol::LinCombOp<aol::MultiVector<RType>, aol::Scalar<RType> > E;
      aol::LinCombOp<aol::MultiVector<RType> > DE;
      if (model == 0) {
        E.appendReference(nccE);
        DE.appendReference(nccDE);
      } else {
        E.appendReference(ssdE);
        DE.appendReference(ssdF);
      }
      E.appendReference(regE, lambda);
      DE.appendReference(regDE, lambda);
      TracingGD gd(grid, E, DE, maxIt, 1.0, stopEps);
      gd.setConfigurationFlags(GDType::USE_NONLINEAR_CG | GDType::DO_NOT_WRITE_CONSOLE_OUTPUT);
      std::stringstream energyLog;
      energyLog.precision(17);
      gd.setOutStream(energyLog);
      gd.applySingle(phi);
      append(out, phi[0]); append(out, phi[1]);
      // Parse the energy log: lines "k-1 [lambda*Ereg] Etotal".
      std::vector<double> etot;
      std::string line;
      while (std::getline(energyLog, line)) {
        std::istringstream ls(line);
        std::vector<double> vals; double v;
        while (ls >> v) vals.push_back(v);
        if (!vals.empty()) etot.push_back(vals.back());
      }
      const int iters = static_cast<int>(etot.size()) - 1;
      out.push_back(iters);
      out.push_back(gd.getEnergyAtLastPosition());
      for (int k = 0; k <= iters; ++k) out.push_back(etot[k]);
      out.push_back(static_cast<double>(gd.taus.size()));
      for (size_t k = 0; k < gd.taus.size(); ++k) out.push_back(gd.taus[k]);
      for (size_t k = 0; k < gd.sigmas.size(); ++k) out.push_back(gd.sigmas[k]);
      for (size_t k = 0; k < gd.evalsPerIter.size(); ++k) out.push_back(gd.evalsPerIter[k]);
    } else if (cmd == "xshift" || cmd == "xshift_solve") {
      // params: lambda yRegFactor [maxIt] ; in: R, T, phix
      // xshift: out E_data, E_reg (1/2 phix.A phix), DE (N^2) of findXShift's E and DE
      // xshift_solve: findXShift's descent from phix (laplaceWeight 0); out phix, iterations, E_final,
      // E_total per iteration (from 0), number of Armijo calls, tau / sigma / evals per call
      const double lambda = atof(argv[3]), yreg = atof(argv[4]);
      const int maxIt = cmd == "xshift_solve" ? atoi(argv[5]) : 0;
      qc::ScalarArray<RType, qc::QC_2D> R(grid), T(grid), phix(grid);
      toVec(&in[0], R); toVec(&in[NN], T); toVec(&in[2 * NN], phix);
      LineEnergy eData(grid, R, T);
      const aol::DerivativeWrapper<RType, LineEnergy, aol::Vector<RType> > deData(eData);
      const aol::FEOpMixedDerivative<ConfType> stiffX(grid, 0, 0, aol::ONTHEFLY);
      const aol::FEOpMixedDerivative<ConfType> stiffY(grid, 1, 1, aol::ONTHEFLY);
      ConfType::MatrixType mat(grid);
      stiffX.assembleAddMatrix(mat);
      stiffY.assembleAddMatrix(mat, yreg);
      const aol::QuadraticFormOp<aol::Vector<RType> > eReg(mat);
      aol::LinCombOp<aol::Vector<RType>, aol::Scalar<RType> > E;
      E.appendReference(eData);
      E.appendReference(eReg, lambda);
      aol::LinCombOp<aol::Vector<RType> > DE;
      DE.appendReference(deData);
      DE.appendReference(mat, lambda);
      if (cmd == "xshift") {
        aol::Scalar<RType> ed, er;
        eData.apply(phix, ed);
        eReg.apply(phix, er);
        out.push_back(ed[0]);
        out.push_back(er[0]);
        aol::Vector<RType> de(NN);
        DE.apply(phix, de);
        append(out, de);
      } else {
        TracingGDV gd(grid, E, DE, maxIt);
        gd.setConfigurationFlags(GDVType::USE_NONLINEAR_CG | GDVType::DO_NOT_WRITE_CONSOLE_OUTPUT);
        std::stringstream energyLog;
        energyLog.precision(17);
        gd.setOutStream(energyLog);
        gd.applySingle(phix);
        append(out, phix);
        std::vector<double> etot;
        std::string line;
        while (std::getline(energyLog, line)) {
          std::istringstream ls(line);
          std::vector<double> vals; double v;
          while (ls >> v) vals.push_back(v);
          if (!vals.empty()) etot.push_back(vals.back());
        }
        const int iters = static_cast<int>(etot.size()) - 1;
        out.push_back(iters);
        out.push_back(gd.getEnergyAtLastPosition());
        for (int k = 0; k <= iters; ++k) out.push_back(etot[k]);
        out.push_back(static_cast<double>(gd.taus.size()));
        for (size_t k = 0; k < gd.taus.size(); ++k) out.push_back(gd.taus[k]);
        for (size_t k = 0; k < gd.sigmas.size(); ++k) out.push_back(gd.sigmas[k]);
        for (size_t k = 0; k < gd.evalsPerIter.size(); ++k) out.push_back(gd.evalsPerIter[k]);
      }
    } else {
      fprintf(stderr, "unknown command %s\n", cmd.c_str());
      return 2;
    }
    writeAll(outPath, out);
  } catch (aol::Exception &e) {
    e.dump();
    return 1;
  }
  return 0;
}
