// Host-side view of the compiled tile-kernel variants (fft_variant.hip, fft_kernels.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include "fft_args.h"

// Fusion modes compiled per variant, in the order fft_variant.hip fills them.
enum sfh_mode_index
{
    SFH_MI_PLAIN = 0,     // FFT only
    SFH_MI_T4 = 1,        // FFT + four-step twiddle (pass A of a two-pass axis)
    SFH_MI_BLUE_FUSED = 2,// chirp, FFT, spectrum, FFT, chirp (Bluestein in one launch)
    SFH_MI_BLUE_A = 3,    // chirp + zero pad, column FFT, twiddle
    SFH_MI_BLUE_B = 4,    // row FFT, spectrum, row FFT, twiddle
    SFH_MI_BLUE_C = 5,    // column FFT, chirp, truncate
    SFH_NUM_MODES = 6
};

struct sfh_kernel_variant
{
    const void *fn[SFH_NUM_MODES]; // __global__ entries per fusion mode
    int nt;                        // threads per workgroup
    int e;                         // elements per thread; LDS capacity = nt * e complex values
    int maxr;                      // largest radix compiled into the variant (16, or 9 when e > 16)
};

extern "C" const sfh_kernel_variant *sfh_variants(int is_double, int *count);
extern "C" int sfh_launch_tile(
    const sfh_kernel_variant *variant, int mode_index, const sfh_tile_args *args,
    hipStream_t stream);
extern "C" int sfh_launch_copy_scale(
    const void *in, void *out, uint64_t count, int is_double, double scale, hipStream_t stream);
