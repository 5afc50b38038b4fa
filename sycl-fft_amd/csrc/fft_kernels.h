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
    SFH_MI_PLAIN_PF = 6,  // persistent, prefetching form of SFH_MI_PLAIN
    SFH_MI_T4_PF = 7,     // persistent, prefetching form of SFH_MI_T4
    SFH_MI_PLAIN_BF16 = 8,// SFH_MI_PLAIN with radix-16 passes on MFMA, split BF16 operands
    SFH_MI_T4_BF16 = 9,   // SFH_MI_T4, same
    SFH_MI_PLAIN_F16 = 10,// SFH_MI_PLAIN with radix-16 passes on MFMA, scaled split FP16
    SFH_MI_T4_F16 = 11,   // SFH_MI_T4, same
    SFH_MI_QUEUE = 12,    // work-queue form of SFH_MI_PLAIN (sfh_queue_args, sfh_launch_queue)
    SFH_NUM_MODES = 13
};

struct sfh_kernel_variant
{
    const void *fn[SFH_NUM_MODES]; // __global__ entries per fusion mode (null: not compiled)
    int nt;                        // threads per workgroup
    int e;                         // elements per thread; LDS capacity = nt * e complex values
    int maxr;                      // largest radix compiled into the variant (16, or 9 when e > 16)
    int tw_global;                 // 1: twiddle tables read from global memory, none in LDS
};

// Length-specialized tile kernels (fft_fixed.hip): one transform of length L per workgroup
// of nt work-items, radix program r1 (and r2 for the Bluestein mode's second FFT) fixed at
// compile time; arguments as for the generic entry of `mode_index`, except twtab: the
// per-pass twiddle tables of the program ([t - 1][j] = W_{ns r}^(j t) for passes 1.. of r1,
// then of r2), and lds_bytes = the entry's own.
// mode_index of the length-specialized Bluestein form over two workgroups per transform
// (no generic counterpart)
#define SFH_FX_BLUE_PAIR 64
// mode_index of the length-specialized C2R kernels (tangle fused into the first pass)
#define SFH_FX_C2R 65
// mode_index of the length-specialized R2C kernels (untangle fused into the last pass)
#define SFH_FX_R2C 66

struct sfh_fixed_kernel
{
    const void *fn;
    int is_double;
    int mode_index; // SFH_MI_PLAIN, SFH_MI_BLUE_FUSED, SFH_FX_BLUE_PAIR, SFH_FX_C2R or SFH_FX_R2C
    int nt;
    uint32_t L;
    uint32_t lds_bytes;
    // used only for launches of at most this many tiles per CU (0: any): entries that spread
    // one transform over more work-items, for batches that give every CU one tile or less
    int max_tiles_per_cu;
    int n1, n2;
    int r1[8], r2[8];
    // columns per tile: 1 (one contiguous transform), or W columns of a strided axis (lanes
    // column-fastest, SFH_WFAST; the last tile of a row may hold fewer columns)
    int W;
    // 1: used only for the half-length FFT of a real plan (where its C2R / R2C forms fuse
    // the tangle or untangle pass), not for complex plans of its length
    int half_only;
    // column tiles: log2 of the group of adjacent tiles run on one XCD (SFH_TILE_XCD_SHIFT)
    int xcd_lg;
};
extern "C" const sfh_fixed_kernel *sfh_fixed_kernels(int *count);
extern "C" int sfh_launch_fixed(const sfh_fixed_kernel *kernel, const sfh_tile_args *args, hipStream_t stream);

extern "C" int sfh_device_cus();

extern "C" const sfh_kernel_variant *sfh_variants(int is_double, int *count);
// grid = workgroups launched: args->ntiles for one-tile-per-workgroup entries, at most the
// resident capacity for the persistent (prefetching) entries
extern "C" int sfh_launch_tile(
    const sfh_kernel_variant *variant, int mode_index, const sfh_tile_args *args, uint32_t grid,
    hipStream_t stream);
// workgroups of one entry resident per CU at the given dynamic LDS (0 if unknown)
extern "C" int sfh_resident_groups(const sfh_kernel_variant *variant, int mode_index, uint32_t lds_bytes);
// one work-queue launch (SFH_MI_QUEUE entry) of `grid` workgroups with `lds_bytes` of dynamic LDS
extern "C" int sfh_launch_queue(
    const sfh_kernel_variant *variant, const sfh_queue_args *args, uint32_t grid, uint32_t lds_bytes,
    hipStream_t stream);
// real transforms: R2C untangle (c2r = 0: M complex -> M + 1 per transform) or C2R tangle
// (c2r = 1: M + 1 -> M), tw = W_N^k, k = 0..M, N = 2M
extern "C" int sfh_launch_real_pass(
    int c2r, const void *in, void *out, const void *tw, uint64_t M, uint64_t batch, int is_double, double scale,
    hipStream_t stream);
// strided / planar layout conversion of batch x n complex values (element x of transform b
// at b * dist + x * stride; planar = separate real and imaginary arrays)
extern "C" int sfh_launch_layout_copy(
    const void *s_re, const void *s_im, int64_t s_stride, int64_t s_dist, int s_planar, void *d_re, void *d_im,
    int64_t d_stride, int64_t d_dist, int d_planar, uint64_t n, uint64_t batch, int is_double, hipStream_t stream);
extern "C" int sfh_launch_copy_scale(
    const void *in, void *out, uint64_t count, int is_double, double scale, hipStream_t stream);
