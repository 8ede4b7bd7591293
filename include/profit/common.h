/*
 * profit/common.h -- common types of the C++ host API
 * (libprofit: include/profit/common.h:53,99-130).
 */
#ifndef PROFIT_AMD_COMMON_H
#define PROFIT_AMD_COMMON_H

#include <chrono>
#include <ostream>

namespace profit {

/* nanoseconds, as libprofit's statistics report them */
typedef std::chrono::nanoseconds::rep nsecs_t;

/* SIMD instruction sets libprofit's BRUTE convolver chooses from; accepted
 * and without effect here (convolution runs on the device) */
enum simd_instruction_set {
	AUTO = 0,
	NONE,
	SSE2,
	AVX
};

template <typename T, typename CharT>
std::basic_ostream<T, CharT> &operator<<(std::basic_ostream<T, CharT> &os, simd_instruction_set instruction_set)
{
	switch (instruction_set) {
	case AUTO: return os << "AUTO";
	case NONE: return os << "NONE";
	case SSE2: return os << "SSE2";
	case AVX: return os << "AVX";
	}
	return os << "UNKNOWN";
}

}  // namespace profit

#endif
