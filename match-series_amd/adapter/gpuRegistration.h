// A registration object for the reference's own series driver whose gradient flow runs on the GPU.
//
// SeriesMatching<RegisType> (projects/electronMicroscopy/matchSeries.h:17-800) is a template over its
// registration object, so a reference build switches to the GPU by naming this class as RegisType
// in matchSeries.cpp (matchSeries.cpp:17-21 for model 0, the "case 22" branch for SSD):
//
//   typedef msg::GpuRegistrationMultilevelDescent<ConfType, qc::SSDRegistrationConfigurator<ConfType> > RegisType;
//   matchSeries<ConfType, RegisType>( parser, argc, argv );
//
// and linking libmsgpu.so. The class is qc::StandardRegistrationMultilevelDescent
// (quoc/registration.h:1072-1132) with one member replaced: descentOnCurrentGrid, the level's
// gradient flow (StandardRegistration::findTransformation, registration.h:967-1000), which is where
// the series spends its time. Everything else SeriesMatching calls (image loading and
// normalisation, the pyramids, prolongation, composition, warping, file output) stays the
// reference's own code. The level solve works on a msg_ctx holding the pair at every level:
//   * the prepared finest images are uploaded when they change (loadRefOrTemplate, setReference,
//     setTemplate) and restricted on the device as MultilevelArray::levRestrict does;
//   * the level's deformation goes up before and comes back after the solve, so the reference's
//     own prolongation and saving see the GPU result;
//   * energy_<LL>.txt is written from the solve's per-iteration record in GradientDescentBase's
//     format ("k lambda*E_reg E", NCC with E - lambda*E_reg first).
// Parameters: "lambda", "lambdaFactor", "maxGDIterations" (default 1000), "stopEpsilon" (default 0)
// as the reference reads them; "gpuFieldPrecision 32" keeps the fields in FP32 (default 64: FP64,
// the reference's own arithmetic). Errors of the library are thrown as aol::Exception.
//
// Builds against the reference's headers only (C++11); oracle/Makefile's "adapter" target compiles a
// reference matchSeries with it as a build check.
#ifndef MSG_GPU_REGISTRATION_H
#define MSG_GPU_REGISTRATION_H

#include <registration.h>

#include <fstream>
#include <vector>

#include "ms_gpu.h"

namespace msg {

// The data term of a reference registration configurator.
template <typename RegistrationConfiguratorType>
struct DataTermOf;
template <typename C>
struct DataTermOf<qc::SSDRegistrationConfigurator<C> > {
  static const int value = MSG_SSD;
};
template <typename C>
struct DataTermOf<qc::NCCRegistrationConfigurator<C> > {
  static const int value = MSG_NCC;
};

inline void check(int rc, const char *call, const char *file, int line) {
  if (rc != MSG_OK) throw aol::Exception(aol::strprintf("%s: %s", call, msg_last_error()).c_str(), file, line);
}
#define MSG_ADAPTER_CHECK(x) ::msg::check((x), #x, __FILE__, __LINE__)

template <typename ConfiguratorType, typename RegistrationConfiguratorType>
class GpuRegistrationMultilevelDescent
    : public qc::StandardRegistrationMultilevelDescent<ConfiguratorType, RegistrationConfiguratorType> {
  typedef qc::StandardRegistrationMultilevelDescent<ConfiguratorType, RegistrationConfiguratorType> Base;

 public:
  typedef typename ConfiguratorType::RealType RealType;
  typedef typename ConfiguratorType::ArrayType ArrayType;

  explicit GpuRegistrationMultilevelDescent(const aol::ParameterParser &Parser)
      : Base(Parser), _ctx(NULL), _imagesOnDevice(false) {
    const int dtype = Parser.getIntOrDefault("gpuFieldPrecision", 64) == 32 ? MSG_F32 : MSG_F64;
    MSG_ADAPTER_CHECK(msg_ctx_create(&_ctx, this->getMaxGridDepth(), 1, dtype));
    MSG_ADAPTER_CHECK(msg_ctx_set_data_term(_ctx, DataTermOf<RegistrationConfiguratorType>::value));
  }

  virtual ~GpuRegistrationMultilevelDescent() { msg_ctx_destroy(_ctx); }

  // SeriesMatching calls these on the RegisType itself: they hide the base class's members.
  void loadRefOrTemplate(const char *Filename, const qc::REGISTRATION_INPUT_TYPE Input, const bool NoScaling) {
    Base::loadRefOrTemplate(Filename, Input, NoScaling);
    _imagesOnDevice = false;
  }
  void setReference(const ArrayType &Reference) {
    Base::setReference(Reference);
    _imagesOnDevice = false;
  }
  void setTemplate(const ArrayType &Template) {
    Base::setTemplate(Template);
    _imagesOnDevice = false;
  }

  // StandardRegistrationMultilevelDescent::descentOnCurrentGrid (registration.h:1092-1112) with the
  // gradient flow on the GPU.
  void descentOnCurrentGrid() {
    const aol::ParameterParser &P = this->getParserReference();
    const int level = this->_curLevel;
    const int n = 1 << level;
    const size_t nn = (size_t)n * n;
    qc::MultiArray<RealType, ConfiguratorType::Dim> phi(this->_transformation, aol::FLAT_COPY);
    const RegistrationConfiguratorType regisConfig(*this);
    const RealType lambda = this->getRegularizationWeight(level);
    this->logLevelStart(lambda);
    if (this->_disableSaving == false) regisConfig.writeInitialRegistration(*this, phi);

    if (!_imagesOnDevice) {
      MSG_ADAPTER_CHECK(msg_set_reference(_ctx, 0, this->_org_reference[this->_maxDepth].getData()));
      MSG_ADAPTER_CHECK(msg_set_template(_ctx, 0, this->_org_template[this->_maxDepth].getData()));
      _imagesOnDevice = true;
    }
    std::vector<double> field(2 * nn);
    for (int c = 0; c < 2; ++c)
      for (size_t i = 0; i < nn; ++i) field[c * nn + i] = phi[c][i];
    MSG_ADAPTER_CHECK(msg_set_deformation_level(_ctx, 0, level, field.data()));

    const int maxIter = P.getIntOrDefault("maxGDIterations", 1000);
    const double stopEps = P.getDoubleOrDefault("stopEpsilon", 0.0);
    double e0 = 0, r0 = 0;
    startEnergy(level, e0, r0);
    std::vector<msg_iter_trace> trace(maxIter > 0 ? maxIter : 1);
    double energy = 0;
    int iterations = 0;
    MSG_ADAPTER_CHECK(msg_solve_level(_ctx, level, lambda, maxIter, stopEps, 0, &energy, &iterations, trace.data(),
                                      maxIter));
    MSG_ADAPTER_CHECK(msg_get_deformation_level(_ctx, 0, level, field.data()));
    for (int c = 0; c < 2; ++c)
      for (size_t i = 0; i < nn; ++i) phi[c][i] = field[c * nn + i];
    this->_energyOfLastSolution = energy;

    if (this->_disableSaving == false) {
      std::ofstream out(aol::strprintf("%senergy_%02d.txt", this->getSaveDirectory(), level).c_str());
      const bool ncc = DataTermOf<RegistrationConfiguratorType>::value == MSG_NCC;
      writeEnergyLine(out, -1, lambda * r0, e0 + lambda * r0, ncc);
      for (int k = 0; k < iterations && k < maxIter; ++k)
        writeEnergyLine(out, k, lambda * trace[k].ereg, trace[k].energy, ncc);
      regisConfig.writeCurrentRegistration(*this, phi);
    }
  }

 private:
  // The level's data and regularisation energies at the deformation just uploaded.
  void startEnergy(int level, double &data, double &reg) {
    void *R = NULL, *T = NULL, *F = NULL;
    MSG_ADAPTER_CHECK(msg_ctx_buffer(_ctx, 0, level, &R));
    MSG_ADAPTER_CHECK(msg_ctx_buffer(_ctx, 1, level, &T));
    MSG_ADAPTER_CHECK(msg_ctx_buffer(_ctx, 2, level, &F));
    const int dtype = this->getParserReference().getIntOrDefault("gpuFieldPrecision", 64) == 32 ? MSG_F32 : MSG_F64;
    if (DataTermOf<RegistrationConfiguratorType>::value == MSG_NCC)
      MSG_ADAPTER_CHECK(msg_ncc_energy(dtype, R, T, F, 1 << level, 1, &data, &reg, NULL));
    else
      MSG_ADAPTER_CHECK(msg_energy(dtype, R, T, F, 1 << level, 1, &data, &reg, NULL));
  }

  static void writeEnergyLine(std::ofstream &out, int k, double regw, double E, bool ncc) {
    out << k << " ";
    if (ncc) out << E - regw << " ";
    out << regw << " " << E << "\n";
  }

  msg_ctx *_ctx;
  bool _imagesOnDevice;
};

}  // namespace msg

#endif  // MSG_GPU_REGISTRATION_H
