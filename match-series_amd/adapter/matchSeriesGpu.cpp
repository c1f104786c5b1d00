// The reference's matchSeries program with its registration object replaced by
// msg::GpuRegistrationMultilevelDescent (deformationModel 22: SSD, 0: NCC, the two registration
// configurators this build has): the reference's own series driver, image handling and file output
// around the GPU level solves. Same grid type as the reference program (matchSeries.cpp:9).
// Built by oracle/Makefile's "adapter" target from the reference's sources where they lie.
#include <matchSeries.h>
#include <cellCenteredGrid.h>

#include "gpuRegistration.h"

typedef qc::RectangularGridConfigurator<double, qc::QC_2D, aol::GaussQuadrature<double, qc::QC_2D, 3>,
                                        qc::CellCenteredCubicGrid<qc::QC_2D> > ConfType;

int main(int argc, char **argv) {
  try {
    aol::ParameterParser parser(argc > 1 ? argv[1] : "matchSeries.par");
    switch (parser.getIntOrDefault("deformationModel", 0)) {
      case 0:
        matchSeries<ConfType, msg::GpuRegistrationMultilevelDescent<ConfType, qc::NCCRegistrationConfigurator<ConfType> > >(
            parser, argc, argv);
        break;
      case 22:
        matchSeries<ConfType, msg::GpuRegistrationMultilevelDescent<ConfType, qc::SSDRegistrationConfigurator<ConfType> > >(
            parser, argc, argv);
        break;
      default:
        throw aol::Exception("matchSeriesGpu: deformationModel 0 (NCC) or 22 (SSD)", __FILE__, __LINE__);
    }
  } catch (aol::Exception &el) {
    el.dump();
  }
  return 0;
}
