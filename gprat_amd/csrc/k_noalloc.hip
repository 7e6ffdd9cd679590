// Fixed device scratch for the entry point that allocates nothing
// (gprat_model_predict_mean_noalloc): a file-scope device array, one copy
// per GPU, whose device address is looked up once per device when the first
// context on that device is created.
//
// In host code the array's name is only a host-side stand-in for the
// symbol; kernels and copies use the address hipGetSymbolAddress returns
// for the current device.
#include <mutex>

#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

constexpr std::size_t kScratchBytes = std::size_t(8) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

std::size_t noalloc_scratch_bytes() { return kScratchBytes; }

void noalloc_scratch_init(int device)
{
    require(device >= 0 && device < kMaxDevices, "no-allocation scratch: device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device])
        return;
    int cur = -1;
    GP_HIP(hipGetDevice(&cur));
    GP_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    GP_HIP(hipSetDevice(cur));
    GP_HIP(e);
    require(p != nullptr, "no-allocation scratch: no device address");
    g_ptr[device] = p;
}

void *noalloc_scratch(int device)
{
    require(device >= 0 && device < kMaxDevices, "no-allocation scratch: device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device])
        throw Error(GPRAT_ERR_INTERNAL, "no-allocation scratch: not set up for device " + std::to_string(device));
    return g_ptr[device];
}

}  // namespace gpamd::k
