// Readers for the image formats that qc::ScalarArray<QC_2D>::load picks by file-name suffix besides
// the netpbm-style arrays (quoc/scalarArray.cpp:547-569): PNG, TIFF, Gatan DM3/DM4 and MetaImage.
// Each returns the values the reference's loader stores in its array of doubles, and fails with an
// exception where the reference's loader throws.
#pragma once

#include <string>

#include "array_io.h"

namespace mshost {

// PNG: decoded with the reference's libpng transforms (expand palette and low bit depths to 8-bit,
// strip alpha); each pixel is the mean of its remaining channel bytes (grey: the byte; colour:
// (r + g + b) / 3). 16-bit PNGs are rejected like the reference rejects them
// (quoc/pngInterface.cpp:5-85, quoc/scalarArray.cpp:573-588).
Image decodePNG(const std::string &bytes, const std::string &name);

// TIFF (classic and BigTIFF, either byte order): the first directory, one sample per pixel, strips
// uncompressed, PackBits, LZW or Deflate, horizontal or floating-point predictor; samples read as
// int / IEEE float / unsigned of 8-64 bits by SampleFormat (quoc/scalarArray.cpp:590-660).
Image decodeTIFF(const std::string &bytes, const std::string &name);

// Gatan DigitalMicrograph 3 and 4: the last ImageList entry's ImageData.Data array with its
// Dimensions, element types int16/int32/uint16/uint32/float32/uint8 (quoc/dm3Import.h:16-436).
Image decodeDM(const std::string &bytes, const std::string &name);

// qc::DM3Reader::saveQuocDataInDM3Container (quoc/dm3Import.h:139-162): a copy of the DM file dmFile
// named outBase + ".dm3" / ".dm4" whose image data (the last ImageList entry) is replaced by the nx x ny
// array, centred into the file's image size (ScalarArray<QC_2D>::padFrom, scalarArray.h:1110-1133,
// zeros around) and converted to the file's element type (int32, uint16, uint32, float32, or uint8
// clamped to [0, 255]).
void saveInDMContainer(const std::string &dmFile, const double *data, int nx, int ny, const std::string &outBase);

// MetaImage header (.mhd) in the fixed layout the reference writes, MET_FLOAT or MET_USHORT data in
// the raw file it names next to the header (quoc/auxiliary.h:51-125, quoc/scalarArray.cpp:664-674).
Image loadMetaImage(const std::string &fileName);

}  // namespace mshost
