// msgio — array file utility over the reference's formats (cf. the reference's image converters):
//   msgio convert <in> <out>     read any supported array (P2/P5/P7/P8/P9/Pa/Pb, optionally .bz2/.q2bz;
//                                .png, .tif/.tiff, .dm3/.dm4, .mhd by suffix) and write it as P9 (float64), bzip2-compressed when <out> ends in .bz2/.q2bz
//   msgio png <in> <out.png>     8-bit grey preview scaled from the value range (the reference's
//                                average/median/numSamples .png)
//   msgio info <in>              print "W H min max mean"
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <numeric>

#include "array_io.h"
#include "params.h"

int main(int argc, char **argv) {
  try {
    if (argc >= 4 && std::string(argv[1]) == "convert") {
      const mshost::Image im = mshost::loadImage(argv[2]);
      mshost::saveP9(argv[3], im.data.data(), im.nx, im.ny);
      return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "png") {
      const mshost::Image im = mshost::loadImage(argv[2]);
      mshost::savePNGScaled(argv[3], im.data.data(), im.nx, im.ny);
      return 0;
    }
    if (argc >= 3 && std::string(argv[1]) == "info") {
      const mshost::Image im = mshost::loadImage(argv[2]);
      const auto mm = std::minmax_element(im.data.begin(), im.data.end());
      const double mean = std::accumulate(im.data.begin(), im.data.end(), 0.0) / std::max<size_t>(im.data.size(), 1);
      printf("%d %d %.17g %.17g %.17g\n", im.nx, im.ny, *mm.first, *mm.second, mean);
      return 0;
    }
    std::cerr << "usage: msgio convert <in> <out> | msgio png <in> <out.png> | msgio info <in>\n";
    return 2;
  } catch (mshost::Exception &e) {
    e.dump(std::cerr);
    return 1;
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
