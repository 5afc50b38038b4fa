// Tile-kernel instantiations for one (precision, threads, elements-per-thread) variant,
// one translation unit per variant so gfx950 builds run in parallel.
// Built with -DSFH_T=<float|double> -DSFH_NT=<threads> -DSFH_E=<elements per thread>
// -DSFH_NAME=<symbol suffix> -DSFH_SPLIT=<1: also the split-precision MFMA entries>.
#include "fft_tile.hpp"

#define SFH_CAT2(a, b) a##b
#define SFH_CAT(a, b) SFH_CAT2(a, b)

// Fusion modes in the order of sfh_mode_index (fft_kernels.h): the six one-tile-per-workgroup
// entries, the persistent prefetching forms of the plain and four-step modes, the
// split-precision forms of the plain and four-step modes (null where not compiled), then
// the work-queue form of the plain mode.
#define SFH_K(MODE, PREF) reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, MODE, PREF>)
extern "C" void SFH_CAT(sfh_kernels_, SFH_NAME)(const void **out)
{
    out[0] = SFH_K(0, 0);
    out[1] = SFH_K(SFH_M_T4, 0);
    out[2] = SFH_K(SFH_M_PRE | SFH_M_MID | SFH_M_POST, 0);
    out[3] = SFH_K(SFH_M_PRE | SFH_M_T4, 0);
    out[4] = SFH_K(SFH_M_MID | SFH_M_T4, 0);
    out[5] = SFH_K(SFH_M_POST, 0);
    out[6] = SFH_K(0, 1);
    out[7] = SFH_K(SFH_M_T4, 1);
#if SFH_SPLIT
    // fp32 variants with radix-16 passes: split-precision MFMA forms of plain and four-step
    out[8] = SFH_K(SFH_M_SPLIT_BF16, 0);
    out[9] = SFH_K(SFH_M_T4 | SFH_M_SPLIT_BF16, 0);
    out[10] = SFH_K(SFH_M_SPLIT_F16, 0);
    out[11] = SFH_K(SFH_M_T4 | SFH_M_SPLIT_F16, 0);
#else
    out[8] = out[9] = out[10] = out[11] = nullptr;
#endif
    // work-queue form of the plain mode (takes sfh_queue_args)
    out[12] = reinterpret_cast<const void *>(&sfh::tile_fft_queue<SFH_T, SFH_NT, SFH_E, 0>);
#if !defined(__HIP__)
    // CPU emulation: hipLaunchKernel needs to know this entry's argument type
    emu_register_queue_kernel(out[12], [](void **args) {
        sfh::tile_fft_queue<SFH_T, SFH_NT, SFH_E, 0>(*static_cast<const sfh_queue_args *>(args[0]));
    });
#endif
}
