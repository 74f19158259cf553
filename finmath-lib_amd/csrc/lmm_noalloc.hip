// fm_lmm_value_products_noalloc: fm_lmm_value_products for callers that cannot allocate
// device memory (tight pricing loops, harnesses that forbid hipMalloc). The per-call device
// buffers of the pass -- per-wavefront (sum, M2) partials, the lowered product and period
// tables and the reduction output -- live in one file-scope device array instead of
// context-owned buffers, and no context buffer may grow during the call.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "fm_internal.hpp"

namespace {

// Maximum scratch of one call: 2 x 8 B x (products + forwards + 1) per 64 paths of partials
// plus the tables. The 30-swaption strip on 40 forwards fits up to about 3.7M paths.
constexpr size_t kScratchBytes = (size_t)64 << 20;

// Every GPU the library runs on holds its own copy of this array until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_table_mu;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of each device's copy
std::mutex g_call_mu[kMaxDevices];       // one call at a time per device: the calls share the array

struct CurrentDevice {
    int prev = -1;
    explicit CurrentDevice(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) throw fm::FmError(FM_ERR_DEVICE, "hipGetDevice failed");
        if (prev != dev && hipSetDevice(dev) != hipSuccess) throw fm::FmError(FM_ERR_DEVICE, "hipSetDevice failed");
    }
    ~CurrentDevice() { if (prev >= 0) (void)hipSetDevice(prev); }
};

// The address of this device's copy of `scratch`, looked up once per device with that device
// current. In host code the array's name stands for a host-side shadow, not device memory, so
// only this address is ever handed to kernels or copies.
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw fm::FmError(FM_ERR_DEVICE, "device index " + std::to_string(device) + " beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_table_mu);
    if (!g_scratch_ptr[device]) {
        CurrentDevice cd(device);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw fm::FmError(FM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

} // namespace

extern "C" int fm_lmm_value_products_noalloc(fm_lmm* ctx, const fm_product* products, int32_t count, double* values,
                                             double* errors) {
    try {
        if (ctx == nullptr) throw fm::FmError(FM_ERR_INVALID_ARGUMENT, "NULL context");
        const int dev = fm::context_device(ctx);
        void* ptr = scratch_address(dev);
        std::lock_guard<std::mutex> lk(g_call_mu[dev]);
        return fm::value_products_in(ctx, products, count, values, errors, ptr, kScratchBytes);
    } catch (const fm::FmError& e) {
        fm::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        fm::set_last_error(e.what());
        return FM_ERR_INTERNAL;
    } catch (...) {
        fm::set_last_error("unknown error");
        return FM_ERR_INTERNAL;
    }
}
