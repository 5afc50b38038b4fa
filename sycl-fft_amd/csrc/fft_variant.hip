// Tile-kernel instantiations for one (precision, threads, elements-per-thread) variant,
// one translation unit per variant so gfx950 builds run in parallel.
// Built with -DSFH_T=<float|double> -DSFH_NT=<threads> -DSFH_E=<elements per thread>
// -DSFH_NAME=<symbol suffix>.
#include "fft_tile.hpp"

#define SFH_CAT2(a, b) a##b
#define SFH_CAT(a, b) SFH_CAT2(a, b)

// Fusion modes in the order of sfh_mode_index() (fft_kernels.h).
extern "C" void SFH_CAT(sfh_kernels_, SFH_NAME)(const void **out)
{
    out[0] = reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, 0>);
    out[1] = reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, SFH_M_T4>);
    out[2] = reinterpret_cast<const void *>(
        &sfh::tile_fft<SFH_T, SFH_NT, SFH_E, SFH_M_PRE | SFH_M_MID | SFH_M_POST>);
    out[3] = reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, SFH_M_PRE | SFH_M_T4>);
    out[4] = reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, SFH_M_MID | SFH_M_T4>);
    out[5] = reinterpret_cast<const void *>(&sfh::tile_fft<SFH_T, SFH_NT, SFH_E, SFH_M_POST>);
}
