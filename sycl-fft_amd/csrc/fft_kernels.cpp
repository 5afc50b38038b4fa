// Variant table and launcher for the tile kernels instantiated in fft_variant.hip.
#include "fft_kernels.h"

#include <mutex>

#define SFH_DECLARE(name) extern "C" void sfh_kernels_##name(const void **out);
SFH_DECLARE(f32_64_16)
SFH_DECLARE(f32_128_16)
SFH_DECLARE(f32_256_16)
SFH_DECLARE(f32_512_16)
SFH_DECLARE(f32_512_20)
SFH_DECLARE(f32_1024_16)
SFH_DECLARE(f32_1024_20)
SFH_DECLARE(f64_64_16)
SFH_DECLARE(f64_128_16)
SFH_DECLARE(f64_256_16)
SFH_DECLARE(f64_512_16)
SFH_DECLARE(f64_512_20)
SFH_DECLARE(f64_256_16g)
SFH_DECLARE(f64_512_16g)
#undef SFH_DECLARE

namespace
{
struct variant_spec
{
    void (*fill)(const void **);
    int nt;
    int e;
    int tw_global; // twin of the (nt, e) variant with its twiddle tables in global memory
};

const variant_spec k_f32[] = {
    {sfh_kernels_f32_64_16, 64, 16},    {sfh_kernels_f32_128_16, 128, 16},
    {sfh_kernels_f32_256_16, 256, 16},  {sfh_kernels_f32_512_16, 512, 16},
    {sfh_kernels_f32_512_20, 512, 20},  {sfh_kernels_f32_1024_16, 1024, 16},
    {sfh_kernels_f32_1024_20, 1024, 20},
};
const variant_spec k_f64[] = {
    {sfh_kernels_f64_64_16, 64, 16},   {sfh_kernels_f64_128_16, 128, 16},
    {sfh_kernels_f64_256_16, 256, 16}, {sfh_kernels_f64_512_16, 512, 16},
    {sfh_kernels_f64_512_20, 512, 20},
    // 4 waves per SIMD and no tables in LDS: f64 tiles of 2049..2560 points (40 KiB) run four
    // workgroups per CU instead of three, and of 4097..5120 points (80 KiB) two instead of
    // one; chosen per launch by resident-workgroup count
    {sfh_kernels_f64_256_16g, 256, 16, 1},
    {sfh_kernels_f64_512_16g, 512, 16, 1},
};

constexpr int k_nf32 = sizeof(k_f32) / sizeof(k_f32[0]);
constexpr int k_nf64 = sizeof(k_f64) / sizeof(k_f64[0]);
sfh_kernel_variant g_f32[k_nf32];
sfh_kernel_variant g_f64[k_nf64];
std::once_flag g_once;

// Allow every entry the whole 160 KiB of dynamic LDS a gfx950 workgroup can hold.
void allow_lds(sfh_kernel_variant &v)
{
    for (const void *fn : v.fn) {
        if (fn) {
            (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        }
    }
    (void)hipGetLastError();
}

void init_tables()
{
    for (int i = 0; i < k_nf32; ++i) {
        k_f32[i].fill(g_f32[i].fn);
        g_f32[i].nt = k_f32[i].nt;
        g_f32[i].e = k_f32[i].e;
        g_f32[i].maxr = k_f32[i].e <= 16 ? 16 : 9;
        g_f32[i].tw_global = k_f32[i].tw_global;
        allow_lds(g_f32[i]);
    }
    for (int i = 0; i < k_nf64; ++i) {
        k_f64[i].fill(g_f64[i].fn);
        g_f64[i].nt = k_f64[i].nt;
        g_f64[i].e = k_f64[i].e;
        g_f64[i].maxr = k_f64[i].e <= 16 ? 16 : 9;
        g_f64[i].tw_global = k_f64[i].tw_global;
        allow_lds(g_f64[i]);
    }
}
} // namespace

extern "C" const sfh_kernel_variant *sfh_variants(int is_double, int *count)
{
    std::call_once(g_once, init_tables);
    if (is_double) {
        *count = k_nf64;
        return g_f64;
    }
    *count = k_nf32;
    return g_f32;
}

extern "C" int sfh_device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess
        || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cus;
}

extern "C" int sfh_resident_groups(const sfh_kernel_variant *variant, int mode_index, uint32_t lds_bytes)
{
    int n = 0;
    if (!variant->fn[mode_index]) {
        return 0;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, variant->fn[mode_index], variant->nt, lds_bytes)
        != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

extern "C" int sfh_launch_tile(
    const sfh_kernel_variant *variant, int mode_index, const sfh_tile_args *args, uint32_t grid,
    hipStream_t stream)
{
    if (args->ntiles == 0) {
        return 0;
    }
    // AQL dispatches carry the grid size in work-items as a 32-bit value
    if (grid == 0 || grid > args->ntiles || static_cast<uint64_t>(grid) * variant->nt > 0xffffffffull
        || mode_index < 0 || mode_index >= SFH_NUM_MODES || mode_index == SFH_MI_QUEUE
        || !variant->fn[mode_index]) {
        return static_cast<int>(hipErrorInvalidConfiguration);
    }
    void *kernel_args[] = {const_cast<sfh_tile_args *>(args)};
    const hipError_t st = hipLaunchKernel(
        variant->fn[mode_index], dim3(grid), dim3(static_cast<uint32_t>(variant->nt)), kernel_args,
        args->lds_bytes, stream);
    return static_cast<int>(st);
}

extern "C" int sfh_launch_queue(
    const sfh_kernel_variant *variant, const sfh_queue_args *args, uint32_t grid, uint32_t lds_bytes,
    hipStream_t stream)
{
    if (args->nitems == 0) {
        return 0;
    }
    if (grid == 0 || grid > args->nitems || !variant->fn[SFH_MI_QUEUE] || lds_bytes > 163840
        || args->lds_next + sizeof(uint32_t) > lds_bytes) {
        return static_cast<int>(hipErrorInvalidConfiguration);
    }
    void *kernel_args[] = {const_cast<sfh_queue_args *>(args)};
    return static_cast<int>(hipLaunchKernel(variant->fn[SFH_MI_QUEUE], dim3(grid),
                                            dim3(static_cast<uint32_t>(variant->nt)), kernel_args, lds_bytes,
                                            stream));
}

extern "C" int sfh_launch_fixed(const sfh_fixed_kernel *kernel, const sfh_tile_args *args, hipStream_t stream)
{
    if (args->ntiles == 0) {
        return 0;
    }
    const uint32_t kw = static_cast<uint32_t>(kernel->W);
    if (args->ntiles > 0xffffffffull / static_cast<uint64_t>(kernel->nt) || args->W != kw || args->L != kernel->L
        || args->lds_bytes > 163840
        || (kw > 1 && ((args->flags & SFH_WFAST) == 0 || (1u << args->logW) != kw
                        || args->nwt != (args->wcount + kw - 1) / kw))) {
        return static_cast<int>(hipErrorInvalidConfiguration);
    }
    static std::once_flag lds_once;
    std::call_once(lds_once, [] {
        int n = 0;
        const sfh_fixed_kernel *k = sfh_fixed_kernels(&n);
        for (int i = 0; i < n; ++i) {
            (void)hipFuncSetAttribute(k[i].fn, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        }
        (void)hipGetLastError();
    });
    void *kernel_args[] = {const_cast<sfh_tile_args *>(args)};
    return static_cast<int>(hipLaunchKernel(kernel->fn, dim3(static_cast<uint32_t>(args->ntiles)),
                                            dim3(static_cast<uint32_t>(kernel->nt)), kernel_args, args->lds_bytes,
                                            stream));
}
