/*
 * profit/library.h -- library-level functions of the C++ host API
 * (libprofit: include/profit/library.h:38-124).
 */
#ifndef PROFIT_AMD_LIBRARY_H
#define PROFIT_AMD_LIBRARY_H

#include <memory>
#include <string>

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

/* The device and arithmetic a Model evaluates with (in the role of
 * libprofit's OpenCLEnvPtr, include/profit/opencl.h:79-168). */
enum class Precision { FP64 = 0, FP32 = 1 };

class GpuEnv {
public:
	GpuEnv(int device, Precision precision);
	~GpuEnv();
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
