/*
 * profit/opencl.h -- libprofit's accelerator handle (include/profit/opencl.h:
 * 40-168) on the HIP device path.  An OpenCLEnvPtr is the device and
 * arithmetic a Model or Convolver evaluates with; the one implementation is
 * GpuEnv (profit/library.h).  There is no OpenCL: the OpenCL timings stay 0
 * and get_version() reports 0.
 */
#ifndef PROFIT_AMD_OPENCL_H
#define PROFIT_AMD_OPENCL_H

#include <map>
#include <memory>
#include <string>

#include "profit/common.h"

namespace profit {

typedef unsigned int cl_ver_t;

struct OpenCL_command_times {
	nsecs_t submit = 0;
	nsecs_t wait = 0;
	nsecs_t exec = 0;
	OpenCL_command_times &operator+=(const OpenCL_command_times &o)
	{
		submit += o.submit;
		wait += o.wait;
		exec += o.exec;
		return *this;
	}
	const OpenCL_command_times operator+(const OpenCL_command_times &o) const
	{
		OpenCL_command_times t = *this;
		return t += o;
	}
};

struct OpenCL_times {
	nsecs_t kernel_prep = 0;
	size_t nwork_items = 0;
	OpenCL_command_times writing_times;
	OpenCL_command_times reading_times;
	OpenCL_command_times filling_times;
	OpenCL_command_times kernel_times;
	nsecs_t total = 0;
};

class OpenCLEnv {
public:
	virtual ~OpenCLEnv() {}
	virtual cl_ver_t get_version() = 0;
	virtual std::string get_platform_name() = 0;
	virtual std::string get_device_name() = 0;
};
typedef std::shared_ptr<OpenCLEnv> OpenCLEnvPtr;

typedef struct _OpenCL_dev_info {
	std::string name;
	cl_ver_t cl_version;
	bool double_support;
} OpenCL_dev_info;

typedef struct _OpenCL_plat_info {
	std::string name;
	cl_ver_t supported_opencl_version;
	std::map<int, OpenCL_dev_info> dev_info;
} OpenCL_plat_info;

/* One platform ("HIP") holding the visible HIP devices (OpenCL version 0,
 * double support). */
std::map<int, OpenCL_plat_info> get_opencl_info();

/* The environment of HIP device `device_idx` (platform 0 only), evaluating
 * in FP64 if use_double, in FP32 bulk arithmetic otherwise;
 * enable_profiling has no effect. */
OpenCLEnvPtr get_opencl_environment(unsigned int platform_idx, unsigned int device_idx, bool use_double,
                                    bool enable_profiling);

}  // namespace profit

#endif
