/*
 * profit/library.h -- library-level functions of the C++ host API
 * (libprofit: include/profit/library.h:38-124).
 */
#ifndef PROFIT_AMD_LIBRARY_H
#define PROFIT_AMD_LIBRARY_H

#include <memory>
#include <string>

#include "profit/common.h"
#include "profit/opencl.h"

namespace profit {

std::string version();
unsigned short version_major();
unsigned short version_minor();
unsigned short version_patch();
std::string version_suffix();

/* Initialises the library (HIP runtime); false on failure, see init_diagnose(). */
bool init();
std::string init_diagnose();
void finish();
std::string finish_diagnose();

/* Feature queries.  The device path replaces OpenMP, FFTW and OpenCL; they
 * report false.  has_gpu() reports whether a HIP device is visible. */
bool has_openmp();
bool has_fftw();
bool has_fftw_with_openmp();
bool has_opencl();
bool has_gpu();
void clear_cache();
/* Whether the host code was built with support for a SIMD instruction set:
 * AUTO and NONE only (convolution runs on the device) */
bool has_simd_instruction_set(simd_instruction_set instruction_set);
/* The OpenCL version libprofit works against: 0 (no OpenCL) */
unsigned short opencl_version_major();
unsigned short opencl_version_minor();

/* The device and arithmetic a Model evaluates with: the implementation of
 * libprofit's OpenCLEnv (include/profit/opencl.h:79-168). */
enum class Precision { FP64 = 0, FP32 = 1 };

class GpuEnv : public OpenCLEnv {
public:
	GpuEnv(int device, Precision precision);
	~GpuEnv() override;
	cl_ver_t get_version() override { return 0; }
	std::string get_platform_name() override { return "HIP"; }
	std::string get_device_name() override;
	GpuEnv(const GpuEnv &) = delete;
	GpuEnv &operator=(const GpuEnv &) = delete;
	int device() const { return device_; }
	Precision precision() const { return precision_; }
	void *handle() const { return ctx_; }  // the C-ABI pk_ctx *

private:
	int device_;
	Precision precision_;
	void *ctx_ = nullptr;
};
typedef std::shared_ptr<GpuEnv> GpuEnvPtr;

/* A shared environment per (device, precision). */
GpuEnvPtr get_gpu_environment(int device = 0, Precision precision = Precision::FP64);

}  // namespace profit

#endif
