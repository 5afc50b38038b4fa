// CPU emulation of the HIP runtime calls the provider makes (TEST INFRASTRUCTURE ONLY).
#include <hip/hip_runtime.h>

#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "fft_args.h"

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
thread_local dim3 blockDim;
thread_local dim3 gridDim;

namespace
{
std::barrier<> *g_bar = nullptr;
std::mutex g_mu;
std::map<uintptr_t, size_t> g_allocs; // base -> size
hipError_t g_last = hipSuccess;
int g_device = 0;
} // namespace

namespace
{
unsigned char *g_dyn_lds = nullptr;
}

unsigned char *emu_dynamic_lds()
{
    return g_dyn_lds;
}

void emu_syncthreads()
{
    g_bar->arrive_and_wait();
}

uint32_t emu_to_bf16(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if (std::isnan(v)) {
        return 0x7fc0u;
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

float emu_from_bf16(uint32_t h)
{
    const uint32_t u = h << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

uint32_t emu_to_f16(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const float a = std::fabs(v);
    if (std::isnan(v)) {
        return sign | 0x7e00u;
    }
    if (a >= 65520.0f) {
        return sign | 0x7c00u;
    }
    if (a < 6.103515625e-05f) { // subnormal: a multiple of 2^-24
        return sign | static_cast<uint32_t>(std::nearbyint(a * 16777216.0f));
    }
    uint32_t au;
    std::memcpy(&au, &a, 4);
    au += 0xfffu + ((au >> 13) & 1u);
    const uint32_t e = (au >> 23) - 127u + 15u;
    return sign | (e << 10) | ((au >> 13) & 0x3ffu);
}

float emu_from_f16(uint32_t h)
{
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) {
        v = std::ldexp(static_cast<float>(m), -24);
    } else if (e == 31) {
        v = m ? NAN : INFINITY;
    } else {
        v = std::ldexp(static_cast<float>(m | 0x400u), static_cast<int>(e) - 25);
    }
    return (h & 0x8000u) ? -v : v;
}

namespace
{
float g_xbuf[1024];
uint32_t g_abuf[1024][4];
uint32_t g_bbuf[1024][4];
} // namespace

float emu_wave_xor(float v, int mask)
{
    const unsigned t = threadIdx.x;
    g_xbuf[t] = v;
    emu_syncthreads();
    const unsigned lane = t & 63u;
    const float r = g_xbuf[(t - lane) + (lane ^ static_cast<unsigned>(mask))];
    emu_syncthreads();
    return r;
}

double emu_wave_xor_d(double v, int mask)
{
    static double buf[1024];
    const unsigned t = threadIdx.x;
    buf[t] = v;
    emu_syncthreads();
    const unsigned lane = t & 63u;
    const double r = buf[(t - lane) + (lane ^ static_cast<unsigned>(mask))];
    emu_syncthreads();
    return r;
}

void emu_mfma_16x16x32(int prec, float *acc, const uint32_t *a, const uint32_t *b)
{
    const unsigned t = threadIdx.x;
    std::memcpy(g_abuf[t], a, 16);
    std::memcpy(g_bbuf[t], b, 16);
    emu_syncthreads();
    const unsigned lane = t & 63u, base = t - lane, n = lane & 15u, g = lane >> 4;
    auto dec = [&](uint32_t w, unsigned half) {
        const uint32_t bits = half ? (w >> 16) : (w & 0xffffu);
        return prec == 1 ? emu_from_bf16(bits) : emu_from_f16(bits);
    };
    for (unsigned j = 0; j < 4; ++j) {
        const unsigned row = 4 * g + j;
        double s = acc[j];
        for (unsigned k = 0; k < 32; ++k) {
            const unsigned kk = k % 8;
            const double av = dec(g_abuf[base + row + 16 * (k / 8)][kk / 2], kk % 2);
            const double bv = dec(g_bbuf[base + n + 16 * (k / 8)][kk / 2], kk % 2);
            s += av * bv;
        }
        acc[j] = static_cast<float>(s);
    }
    emu_syncthreads();
}

void emu_launch(dim3 grid, dim3 block, const std::function<void()> &body)
{
    const unsigned nt = block.x * block.y * block.z;
    std::barrier<> bar(nt);
    g_bar = &bar;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            threadIdx = dim3(t);
            blockDim = block;
            gridDim = grid;
            for (unsigned bz = 0; bz < grid.z; ++bz) {
                for (unsigned by = 0; by < grid.y; ++by) {
                    for (unsigned bx = 0; bx < grid.x; ++bx) {
                        blockIdx = dim3(bx, by, bz);
                        body();
                        bar.arrive_and_wait();
                    }
                }
            }
        });
    }
    for (auto &x : th) {
        x.join();
    }
    g_bar = nullptr;
}

namespace
{
std::mutex g_qmu;
std::map<const void *, void (*)(void **)> g_queue_kernels;
} // namespace

void emu_register_queue_kernel(const void *fn, void (*call)(void **args))
{
    std::lock_guard<std::mutex> l(g_qmu);
    g_queue_kernels[fn] = call;
}

uint32_t emu_atomic_add_u32(uint32_t *p, uint32_t v)
{
    std::lock_guard<std::mutex> l(g_qmu);
    const uint32_t old = *p;
    *p = old + v;
    return old;
}

hipError_t hipLaunchKernel(const void *fn, dim3 grid, dim3 block, void **args, size_t shmem, hipStream_t)
{
    void (*queue_call)(void **) = nullptr;
    {
        std::lock_guard<std::mutex> l(g_qmu);
        const auto it = g_queue_kernels.find(fn);
        if (it != g_queue_kernels.end()) {
            queue_call = it->second;
        }
    }
    if (queue_call) {
        if (block.x == 0 || block.x > 1024 || grid.x == 0 || shmem > 163840) {
            return hipErrorInvalidConfiguration;
        }
        unsigned char *lds = static_cast<unsigned char *>(std::malloc(shmem ? shmem : 1));
        g_dyn_lds = lds;
        emu_launch(grid, block, [&]() { queue_call(args); });
        g_dyn_lds = nullptr;
        std::free(lds);
        return hipSuccess;
    }
    // every kernel launched through this entry takes one sfh_tile_args by value
    auto kernel = reinterpret_cast<void (*)(sfh_tile_args)>(const_cast<void *>(fn));
    const sfh_tile_args a = *static_cast<const sfh_tile_args *>(args[0]);
    if (block.x == 0 || block.x > 1024 || grid.x == 0 || shmem > 163840) {
        return hipErrorInvalidConfiguration;
    }
    // exactly the requested dynamic LDS, so AddressSanitizer flags any access past it
    unsigned char *lds = static_cast<unsigned char *>(std::malloc(shmem ? shmem : 1));
    g_dyn_lds = lds;
    emu_launch(grid, block, [&]() { kernel(a); });
    g_dyn_lds = nullptr;
    std::free(lds);
    return hipSuccess;
}

hipError_t hipGetDeviceCount(int *c)
{
    *c = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = g_device;
    return hipSuccess;
}
hipError_t hipSetDevice(int d)
{
    if (d != 0) {
        return hipErrorInvalidDevice;
    }
    g_device = d;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->name, "cpu-emulated MI355X");
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->multiProcessorCount = 256;
    return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t n)
{
    *p = std::malloc(n ? n : 1);
    if (!*p) {
        return hipErrorOutOfMemory;
    }
    std::lock_guard<std::mutex> l(g_mu);
    g_allocs[reinterpret_cast<uintptr_t>(*p)] = n;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) {
        return hipSuccess;
    }
    std::lock_guard<std::mutex> l(g_mu);
    g_allocs.erase(reinterpret_cast<uintptr_t>(p));
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    std::memcpy(d, s, n); // work runs in issue order, synchronously
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
    *e = reinterpret_cast<hipEvent_t>(new int(0));
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete reinterpret_cast<int *>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t)
{
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t)
{
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned)
{
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t)
{
    return hipSuccess;
}
static bool find_alloc(const void *p, uintptr_t &base, size_t &size)
{
    std::lock_guard<std::mutex> l(g_mu);
    const auto q = reinterpret_cast<uintptr_t>(p);
    auto it = g_allocs.upper_bound(q);
    if (it == g_allocs.begin()) {
        return false;
    }
    --it;
    if (q >= it->first && q < it->first + std::max<size_t>(it->second, 1)) {
        base = it->first;
        size = it->second;
        return true;
    }
    return false;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *p)
{
    uintptr_t b;
    size_t s;
    std::memset(a, 0, sizeof(*a));
    if (!find_alloc(p, b, s)) {
        g_last = hipErrorInvalidValue;
        return hipErrorInvalidValue;
    }
    a->type = hipMemoryTypeDevice;
    a->device = 0;
    a->devicePointer = const_cast<void *>(p);
    return hipSuccess;
}
hipError_t hipMemGetAddressRange(hipDeviceptr_t *base, size_t *size, hipDeviceptr_t p)
{
    uintptr_t b;
    size_t s;
    if (!find_alloc(p, b, s)) {
        return hipErrorInvalidValue;
    }
    *base = reinterpret_cast<void *>(b);
    *size = s;
    return hipSuccess;
}
hipError_t hipGetLastError()
{
    const hipError_t e = g_last;
    g_last = hipSuccess;
    return e;
}
const char *hipGetErrorString(hipError_t e)
{
    return e == hipSuccess ? "success" : "emulated HIP error";
}

hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int)
{
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t)
{
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s)
{
    *s = nullptr;
    return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *st)
{
    *st = hipStreamCaptureStatusNone; // the emulation has no graphs
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = nullptr;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t)
{
    return hipSuccess;
}

// a small emulated device: 2 CUs, 1 resident workgroup each, so persistent launches loop
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int)
{
    *v = 2;
    return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t)
{
    *n = 1;
    return hipSuccess;
}

hipError_t hipMemset(void *d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    return hipEventCreateWithFlags(e, 0);
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 1.0f;
    return hipSuccess;
}
