// Shared definitions for the gfx950 GP backend: status handling, launch
// helpers and the device-side SE covariance entry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "gprat_amd.h"

namespace gpamd
{

// Exception carrying a C-ABI status; caught at the extern "C" boundary.
struct Error : std::runtime_error
{
    int status;
    Error(int s, const std::string &msg) : std::runtime_error(msg), status(s) { }
};

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line);

#define GP_HIP(expr)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t _e = (expr);                                                                                        \
        if (_e != hipSuccess)                                                                                          \
            ::gpamd::throw_hip(_e, #expr, __FILE__, __LINE__);                                                         \
    } while (0)

#define GP_CHECK_LAUNCH() GP_HIP(hipGetLastError())

inline void require(bool ok, const std::string &msg)
{
    if (!ok)
        throw Error(GPRAT_ERR_ARGUMENT, msg);
}

inline long ceil_div(long a, long b) { return (a + b - 1) / b; }
inline long round_up(long a, long b) { return ceil_div(a, b) * b; }

// Hyperparameters in the form the kernels consume.
struct SEK
{
    double lengthscale;
    double vertical;
    double noise;
    // -0.5 / (l * l), evaluated as the reference does
    // (core/src/cpu/gp_algorithms.cpp:27).
    double scale() const { return -0.5 / (lengthscale * lengthscale); }
};

#ifdef __HIPCC__
// Squared distance between the windows x[i..i+r) and z[j..j+r), accumulated
// in k order with separate multiply and add, so the value matches the
// reference's host loop bit for bit.
__device__ __forceinline__ double window_sqdist(const double *__restrict__ x, long i, const double *__restrict__ z, long j, int r)
{
#pragma clang fp contract(off)
    double d = 0.0;
    for (int k = 0; k < r; k++)
    {
        double t = x[i + k] - z[j + k];
        d += t * t;
    }
    return d;
}
#endif

}  // namespace gpamd
