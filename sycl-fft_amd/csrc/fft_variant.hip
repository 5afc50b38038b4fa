// Tile-kernel instantiations for one (precision, threads, elements-per-thread) variant,
// one translation unit per variant so gfx950 builds run in parallel.
// Built with -DSFH_T=<float|double> -DSFH_NT=<threads> -DSFH_E=<elements per thread>
// -DSFH_NAME=<symbol suffix>.
#include "fft_tile.hpp"

#define SFH_CAT2(a, b) a##b
#define SFH_CAT(a, b) SFH_CAT2(a, b)

// Fusion modes in the order of sfh_mode_index (fft_kernels.h): the six one-tile-per-workgroup
// entries, then the persistent prefetching forms of the plain and four-step modes.
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
}
