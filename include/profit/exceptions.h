/*
 * profit/exceptions.h -- exception types of the C++ host API, the same
 * hierarchy as libprofit's (include/profit/exceptions.h:41-105):
 *   exception <- invalid_parameter <- unknown_parameter, gpu_error (in the
 *   role of opencl_error, which stays as an alias), fft_error, fs_error.
 * The C-ABI status codes (profit_hip.h) map onto them one to one.
 */
#ifndef PROFIT_AMD_EXCEPTIONS_H
#define PROFIT_AMD_EXCEPTIONS_H

#include <exception>
#include <string>

namespace profit {

class exception : public std::exception {
public:
	explicit exception(const std::string &what) : m_what(what) {}
	const char *what() const noexcept override { return m_what.c_str(); }

private:
	std::string m_what;
};

class invalid_parameter : public exception {
public:
	explicit invalid_parameter(const std::string &what) : exception(what) {}
};

class unknown_parameter : public invalid_parameter {
public:
	explicit unknown_parameter(const std::string &what) : invalid_parameter(what) {}
};

/* HIP runtime / device failures (libprofit: opencl_error) */
class gpu_error : public exception {
public:
	explicit gpu_error(const std::string &what) : exception(what) {}
};
typedef gpu_error opencl_error;

class fft_error : public exception {
public:
	explicit fft_error(const std::string &what) : exception(what) {}
};

class fs_error : public exception {
public:
	explicit fs_error(const std::string &what) : exception(what) {}
};

/* Throws the exception that corresponds to a non-zero C-ABI status, with
 * pk_last_error() as its message. */
void throw_status(int status);

}  // namespace profit

#endif
