/*
 * profit/fits.h -- minimal FITS image I/O of the C++ host API
 * (libprofit: include/profit/fits_utils.h, src/fits_utils.cpp): one
 * BITPIX = -64 (big-endian FP64) 2-D primary HDU in 2880-byte blocks, with
 * the pixel scale in CDELT1/CDELT2.
 */
#ifndef PROFIT_AMD_FITS_H
#define PROFIT_AMD_FITS_H

#include <string>

#include "profit/exceptions.h"
#include "profit/image.h"

namespace profit {

class invalid_file : public exception {
public:
	explicit invalid_file(const std::string &what) : exception(what) {}
};

/* Reads the image and its pixel scale. */
Image from_fits(const std::string &filename, PixelScale &pixel_scale);
/* Writes the image; ".fits" is appended to the name if missing.  CRVALn
 * record the world coordinate of the first pixel, (0.5 - offset) * scale. */
void to_fits(const Image &image, const Point &offset, const PixelScale &pixel_scale, std::string filename);

}  // namespace profit

#endif
