// The reference's array files: netpbm-style headers "P<t>\n#comment\nW H\nmax\n" followed by raw or
// ASCII data (quoc/scalarArray.cpp:321-399, 481-555), bzip2-compressed when the name ends in .bz2
// or .q2bz (aol/bzipiostream.h). Types: P2 (8-bit ASCII), P5 (8-bit, or 16-bit big-endian when
// max > 255), P7 (float ASCII), P8 (float32), P9 (float64), Pa (uint16), Pb (int16); little-endian
// binary data.
#pragma once

#include <string>
#include <vector>

namespace mshost {

struct Image {
  int nx = 0, ny = 0;
  std::vector<double> data;  // x fastest
};

Image loadImage(const std::string &fileName);
// PGM_DOUBLE_BINARY (P9), the type the reference writes deformations and reconstructions with.
void saveP9(const std::string &fileName, const double *data, int nx, int ny, const char *comment = nullptr);
// 8-bit grey PNG of the array scaled from its value range [min, max] to [0, 255] by truncation, as
// DefaultArraySaver(true, ...) writes the reconstruction previews (quocTimestepSaver.h:417-441 with
// setOverflowHandlingToCurrentValueRange; CLIP_THEN_SCALE in Vector::createOverflowHandledData,
// vec.cpp:240-252; PNGSaveInterface::loadDataFromScalarArray2d, pngInterface.cpp:136-159).
void savePNGScaled(const std::string &fileName, const double *data, int nx, int ny);
// The same with the range [mn, mx] given (values clipped to it first): setOverflowHandling
// (aol::CLIP_THEN_SCALE, mn, mx), e.g. the saturated range of setOverflowHandlingToEnhanceContrast
// (aol/vec.h:342-345).
void savePNGClipScaled(const std::string &fileName, const double *data, int nx, int ny, double mn, double mx);

// Component-wise load/save of a displacement field from a "%d" file-name mask
// (qc::MultiArray::load / save, quoc/multiArray.h:272-306).
void loadField(const std::string &mask, int n, std::vector<double> &phi);
void saveField(const std::string &mask, const double *phi, int n);

bool hasBzipSuffix(const std::string &f);

}  // namespace mshost
