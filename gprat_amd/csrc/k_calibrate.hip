// Calibration kernels: the matrix-core rate and the HBM copy bandwidth this
// device actually reaches, so measured rates can be reported against both
// the vendor peak and a measured one. The MFMA loops run independent
// accumulators back to back from registers (no memory traffic); the copy
// streams 16-byte vectors through every CU. The loops are held to two
// waves per SIMD: left free, the compiler moves every accumulator between
// VGPRs and AGPRs on each iteration.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int CHAINS = 16;  // independent accumulators per wavefront

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma_f64_loop(double *out, int iters)
{
    v4d acc[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; j++)
        acc[j] = v4d{ 0.0, 0.0, 0.0, 0.0 };
    const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * blockIdx.x;
    for (int i = 0; i < iters; i++)
    {
#pragma unroll
        for (int j = 0; j < CHAINS; j++)
            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < CHAINS; j++)
        s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma_f32_loop(double *out, int iters)
{
    v4f acc[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; j++)
        acc[j] = v4f{ 0.0f, 0.0f, 0.0f, 0.0f };
    const float a = 1.0f + 1e-6f * threadIdx.x, b = 1.0f - 1e-6f * blockIdx.x;
    for (int i = 0; i < iters; i++)
    {
#pragma unroll
        for (int j = 0; j < CHAINS; j++)
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < CHAINS; j++)
        s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_copy_f64(const double *__restrict__ src, double *__restrict__ dst, long pairs)
{
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < pairs; i += stride)
        reinterpret_cast<d2 *>(dst)[i] = reinterpret_cast<const d2 *>(src)[i];
}

}  // namespace

long mfma_loop(hipStream_t s, bool fp64, double *out, int workgroups, int iters)
{
    require(workgroups >= 1 && iters >= 1, "mfma_loop: bad shape");
    if (fp64)
        hipLaunchKernelGGL(k_mfma_f64_loop, dim3(workgroups), dim3(256), 0, s, out, iters);
    else
        hipLaunchKernelGGL(k_mfma_f32_loop, dim3(workgroups), dim3(256), 0, s, out, iters);
    GP_CHECK_LAUNCH();
    // four wavefronts per workgroup, CHAINS products of 16 x 16 x 4 per iteration
    return static_cast<long>(workgroups) * 4 * iters * CHAINS * (16L * 16 * 4 * 2);
}

void copy_f64(hipStream_t s, const double *src, double *dst, long count, int workgroups)
{
    require(count % 2 == 0 && workgroups >= 1, "copy_f64: bad shape");
    require(reinterpret_cast<uintptr_t>(src) % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0, "copy_f64: unaligned");
    hipLaunchKernelGGL(k_copy_f64, dim3(workgroups), dim3(256), 0, s, src, dst, count / 2);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
