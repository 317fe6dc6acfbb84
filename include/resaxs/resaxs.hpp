// resaxs.hpp -- core types and errors of the resaxslib C++ surface, served by the
// MI355X engine (libresaxs_amd.so, C ABI in rsx_capi.h).
//
// Source-compatible with resaxslib/include/resaxs.hpp:29-221: the wavefront-size
// macros, real_type<FLT>/real4, device_type, dev_id, error (same codes) and
// verify().  Bodies stay the reference's cl_float4 / cl_double4 records, taken
// from the OpenCL platform header when the toolchain has one and declared with
// the same layout otherwise.
#ifndef RESAXS_AMD_RESAXS_HPP
#define RESAXS_AMD_RESAXS_HPP

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../rsx_capi.h"

#if __has_include(<CL/cl_platform.h>)
#ifndef CL_TARGET_OPENCL_VERSION
#define CL_TARGET_OPENCL_VERSION 120  // the OpenCL 1.2 types the reference's records use
#endif
#include <CL/cl_platform.h>
#else
typedef float cl_float;
typedef double cl_double;
typedef unsigned int cl_uint;
typedef union {
    cl_float s[4];
    struct {
        cl_float x, y, z, w;
    };
} __attribute__((aligned(16))) cl_float4;
typedef union {
    cl_double s[4];
    struct {
        cl_double x, y, z, w;
    };
} __attribute__((aligned(32))) cl_double4;
#endif

// Workgroup sizes accepted by saxs_params::initialize (the engine picks its own
// launch shapes; the value is kept and reported only)
#define SAXS_NV_FERMI_WF_SIZE 64
#define SAXS_AMD_WF_SIZE 64
#define SAXS_AMD_LOW_END_WF_SIZE 32

namespace resaxs {

template <typename FLT_T>
union real_type {
};

template <>
union real_type<float> {
    typedef float real;
    typedef cl_float4 real4;
    typedef cl_float4 cl_real4;
};

template <>
union real_type<double> {
    typedef double real;
    typedef cl_double4 real4;
    typedef cl_double4 cl_real4;
};

template <typename FLT_T>
using real4 = typename real_type<FLT_T>::real4;

// kinds of device a dev_id may ask for; this library runs on HIP GPUs only
enum device_type { CPU, GPU, ANY };

// a device: platform is 0 (the HIP platform) or -1 (any), device a HIP ordinal or -1 (the first)
struct dev_id {
    long platform;
    long device;
    device_type type;

    bool operator==(const dev_id& o) const { return platform == o.platform && device == o.device && type == o.type; }
    bool operator!=(const dev_id& o) const { return !(*this == o); }
    std::string to_string() const { return std::to_string(platform) + '-' + std::to_string(device); }
};

// the reference's exception type; codes equal the C ABI's RSX_* values
class error : public std::exception {
    int err_;
    std::string msg_;

public:
    explicit error(int err) : err_(err) {}
    error(int err, const char* msg) : err_(err), msg_(msg ? msg : "") {}
    ~error() throw() {}
    const char* what() const throw() override { return msg_.empty() ? "empty" : msg_.c_str(); }
    int err() const { return err_; }

    enum code {
        SAXS_INVALID_DEVICE = RSX_INVALID_DEVICE,
        SAXS_ALG_NOT_INITIALIZED = RSX_ALG_NOT_INITIALIZED,
        SAXS_INVALID_ARG = RSX_INVALID_ARG,
        SAXS_PROGRAM_NOT_LOADED = RSX_PROGRAM_NOT_LOADED,
        SAXS_ARGS_NOT_SET = RSX_ARGS_NOT_SET,
        SAXS_NO_INITIAL_CALC = RSX_NO_INITIAL_CALC,
    };
};

inline void verify(bool condition, int err, const char* msg = nullptr)
{
    if (!condition) throw error(err, msg ? msg : "RESAXS::error");
}

inline void verify(bool condition, int err, const std::string& msg)
{
    verify(condition, err, msg.c_str());
}

namespace detail {

// turns a C ABI status into resaxs::error with the engine's message
inline void check(int rc, const rsx_engine* h, const char* what)
{
    if (rc == RSX_OK) return;
    std::string m = what;
    const char* e = h ? rsx_last_error(h) : nullptr;
    if (e && *e) m += std::string(": ") + e;
    throw error(rc, m.c_str());
}

// bodies of either precision as the engine's float4 records (x, y, z, type bits)
inline void to_float4(const std::vector<cl_float4>& in, std::vector<float>& out)
{
    out.resize(4 * in.size());
    if (!in.empty()) std::memcpy(out.data(), in.data(), out.size() * sizeof(float));
}

inline void to_float4(const std::vector<cl_double4>& in, std::vector<float>& out)
{
    out.resize(4 * in.size());
    for (size_t i = 0; i < in.size(); ++i) {
        for (int c = 0; c < 3; ++c) out[4 * i + c] = (float)in[i].s[c];
        unsigned long long t;
        std::memcpy(&t, &in[i].s[3], sizeof t);
        const unsigned int t32 = (unsigned int)t;
        std::memcpy(&out[4 * i + 3], &t32, sizeof t32);
    }
}

}  // namespace detail
}  // namespace resaxs

#endif
