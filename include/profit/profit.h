/*
 * profit/profit.h -- the C++ host API of the gfx950 image-generation path
 * (libprofit's umbrella header include/profit/profit.h).
 */
#ifndef PROFIT_AMD_PROFIT_H
#define PROFIT_AMD_PROFIT_H

#include "profit/common.h"
#include "profit/convolve.h"
#include "profit/exceptions.h"
#include "profit/image.h"
#include "profit/library.h"
#include "profit/model.h"
#include "profit/opencl.h"
#include "profit/profile.h"

#endif
