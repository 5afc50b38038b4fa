// Plan execution that allocates nothing: syclfft_hip_plan_execute_noalloc.
//
// For callers that cannot allocate device memory while they run (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  A plan's multi-launch
// forms (four-step axes, the three-launch Bluestein form, multi-D plans with such axes)
// pass their intermediates through scratch memory; this entry point runs those launches
// with the file-scope device array below instead of the plan's own scratch.  The array is
// k_scratch_bytes per device, reserved when the library's code object is loaded on a
// device and held until the process ends.
//
// Every call on one device shares that device's array, and the array keeps whatever the
// last call left in it (each launch writes every scratch element it later reads, so no
// clearing is needed).  Calls whose executions overlap on one device race on it: the
// caller orders them (one stream, or events).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "syclfft_hip_provider.h"

extern "C" int sfh_plan_execute_with_scratch(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps, void *done,
    void *scratch, uint64_t scratch_bytes, char *err, size_t cap);
extern "C" int sfh_plan_device(const void *plan);

namespace
{
    constexpr size_t k_scratch_bytes = size_t(64) << 20;
    constexpr int k_max_devices = 64;

    // The scratch array.  In host code its name is a host-side stand-in, not a device
    // address: kernels get the per-device address from hipGetSymbolAddress (scratch_on).
    static __device__ __align__(256) unsigned char scratch[k_scratch_bytes];

    struct device_scratch
    {
        std::once_flag once;
        void *ptr{};
        hipError_t status{hipSuccess};
    };
    device_scratch g_scratch[k_max_devices];

    void put_error(char *err, size_t cap, const std::string &msg)
    {
        if (err && cap) {
            err[msg.copy(err, cap - 1)] = '\0';
        }
    }

    // Address of `scratch` on device `dev`, looked up once per device with that device current.
    hipError_t scratch_on(int dev, void **ptr)
    {
        device_scratch &slot = g_scratch[dev];
        std::call_once(slot.once, [&] {
            int prev = -1;
            if ((slot.status = hipGetDevice(&prev)) != hipSuccess) {
                return;
            }
            if ((slot.status = hipSetDevice(dev)) != hipSuccess) {
                return;
            }
            slot.status = hipGetSymbolAddress(&slot.ptr, HIP_SYMBOL(scratch));
            const hipError_t back = hipSetDevice(prev);
            if (slot.status == hipSuccess) {
                slot.status = back;
            }
            if (slot.status != hipSuccess) {
                slot.ptr = nullptr;
            }
        });
        *ptr = slot.ptr;
        return slot.status;
    }
} // namespace

extern "C" uint64_t syclfft_hip_noalloc_scratch_limit(void)
{
    return k_scratch_bytes;
}

extern "C" int syclfft_hip_plan_execute_noalloc(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps, void *done, char *err,
    size_t cap)
{
    put_error(err, cap, "");
    if (!plan) {
        put_error(err, cap, "plan handle is null");
        return SYCLFFT_HIP_STATUS_INVALID_STATE;
    }
    const int dev = sfh_plan_device(plan);
    if (dev < 0 || dev >= k_max_devices) {
        put_error(err, cap, "plan device " + std::to_string(dev) + " is outside 0.." + std::to_string(k_max_devices - 1));
        return SYCLFFT_HIP_STATUS_INVALID_ARGUMENT;
    }
    void *ptr = nullptr;
    const hipError_t st = scratch_on(dev, &ptr);
    if (st != hipSuccess || !ptr) {
        put_error(err, cap, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(st));
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
    return sfh_plan_execute_with_scratch(plan, in, out, stream, deps, ndeps, done, ptr, k_scratch_bytes, err, cap);
}
