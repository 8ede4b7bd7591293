/*
 * profit/profile.h -- profiles of a Model with libprofit's named-parameter
 * interface (include/profit/profile.h:108-197, src/profile.cpp:69-163).
 * Parameters are registered by name with a type (bool, unsigned int,
 * double); Profile::parameter sets them by value or from "name=value".
 */
#ifndef PROFIT_AMD_PROFILE_H
#define PROFIT_AMD_PROFILE_H

#include <functional>
#include <map>
#include <memory>
#include <string>

#include "profit/common.h"
#include "profit/opencl.h"
#include "profit_hip.h"

namespace profit {

class Model;

/* Statistics of the last evaluation (libprofit's ProfileStats,
 * include/profit/profile.h:47-73).  The device evaluates all profiles of a
 * model together: `total` is the model's device time (setup to returned
 * image), as are refined_pixels and subcells the model's counts. */
class ProfileStats {
public:
	virtual ~ProfileStats() {}
	nsecs_t total = 0;
	double total_ms = 0;
	unsigned long long refined_pixels = 0, subcells = 0;
};

/* Sub-pixel integration statistics of radial profiles; subsampling.total
 * holds the model's profile stage (pixel pass + refinement) */
struct radial_subsampling_stats {
	nsecs_t pre_subsampling = 0;
	nsecs_t new_subsampling = 0;
	nsecs_t inital_transform = 0;
	OpenCL_times cl_times;
	nsecs_t final_transform = 0;
	nsecs_t total = 0;
};

struct RadialProfileStats : ProfileStats {
	OpenCL_times cl_times;
	radial_subsampling_stats subsampling;
	nsecs_t final_image = 0;
};

class Profile {
public:
	Profile(const std::string &name, int kind);
	virtual ~Profile() = default;

	/* Profile::parameter(name, value): invalid_parameter if no parameter of
	 * that name and type exists */
	void parameter(const std::string &name, bool value);
	void parameter(const std::string &name, double value);
	void parameter(const std::string &name, unsigned int value);
	void parameter(const std::string &name, int value) { parameter(name, (unsigned int)value); }
	/* "name=value": unknown_parameter if no parameter of that name exists,
	 * invalid_parameter if the value does not parse as its type */
	void parameter(const std::string &spec);

	const std::string &get_name() const { return name; }
	bool do_convolve() const { return component.convolve != 0; }
	std::shared_ptr<ProfileStats> get_stats() const { return stats; }

	/* the C-ABI description of this profile */
	const pk_component &get_component() const { return component; }
	int get_kind() const { return component.kind; }

private:
	friend class Model;
	template <typename T> using holder = std::map<std::string, std::reference_wrapper<T>>;
	void register_bool(const char *n, int32_t &f) { bools.insert({n, std::ref(f)}); }
	void register_uint(const char *n, uint32_t &f) { uints.insert({n, std::ref(f)}); }
	void register_double(const char *n, double &f) { doubles.insert({n, std::ref(f)}); }

	std::string name;
	pk_component component;
	holder<int32_t> bools;
	holder<uint32_t> uints;
	holder<double> doubles;
	std::shared_ptr<ProfileStats> stats;
};

typedef std::shared_ptr<Profile> ProfilePtr;

}  // namespace profit

#endif
