// Length-specialized tile kernels (fft_fixed.hpp) and their registry.  Each entry names
// its precision, fusion mode, length, workgroup size and radix program; the planner
// (fft_plan.cpp) replaces a generic launch of the same arguments by an entry whose
// (precision, mode, length) match, and builds the entry's per-pass twiddle table from
// its radix program.
#include "fft_fixed.hpp"
#include "fft_kernels.h"

namespace
{
using sfh::fx::radices;

template <typename T, int NT, int WAVES, int MODE, bool PAD, bool SPLIT, bool PF, class R1, class R2, int W = 1>
sfh_fixed_kernel entry(int max_tiles_per_cu = 0, int half_only = 0, int xcd_lg = 0)
{
    using prog = sfh::fx::program<R1, R2, MODE, PAD, W>;
    sfh_fixed_kernel k{};
    k.fn = reinterpret_cast<const void *>(&sfh::fixed_fft<T, NT, WAVES, MODE, PAD, SPLIT, PF, R1, R2, W>);
    k.W = W;
    k.is_double = sizeof(T) == 8;
    k.mode_index = (MODE & SFH_M_PAIR)  ? SFH_FX_BLUE_PAIR
                   : (MODE & SFH_M_MID) ? SFH_MI_BLUE_FUSED
                   : (MODE & SFH_M_C2R) ? SFH_FX_C2R
                   : (MODE & SFH_M_R2C) ? SFH_FX_R2C
                                        : SFH_MI_PLAIN;
    k.nt = NT;
    k.L = static_cast<uint32_t>(prog::L);
    // LDS bytes per tile: one complex value per slot, or one real with the split exchange
    // (W columns: W regions of the column pitch)
    constexpr int elem = static_cast<int>(SPLIT ? sizeof(T) : 2 * sizeof(T));
    k.lds_bytes = W > 1 ? static_cast<uint32_t>(W * prog::column_pitch(elem) * elem)
                        : static_cast<uint32_t>(prog::slots() * elem * ((MODE & SFH_M_DBUF) ? 2 : 1));
    k.max_tiles_per_cu = max_tiles_per_cu;
    k.half_only = half_only;
    k.xcd_lg = xcd_lg;
    k.n1 = R1::n;
    for (int i = 0; i < R1::n; ++i) {
        k.r1[i] = R1::v[i];
    }
    k.n2 = (MODE & SFH_M_MID) ? R2::n : 0;
    for (int i = 0; i < k.n2; ++i) {
        k.r2[i] = R2::v[i];
    }
    return k;
}

constexpr int BLUE = SFH_M_PRE | SFH_M_MID | SFH_M_POST;
using none = radices<1>;

const sfh_fixed_kernel *table(int *count)
{
    static const sfh_fixed_kernel k[] = {
        // complex<float> Bluestein convolution of length 20160 = 2^6 3^2 5 7 (n = 5004 .. 10080,
        // e.g. the prime 10007): odd radices first, so that no power-of-two pass writes at a
        // small ns (bank conflicts); 157.5 KiB of LDS, no room for the pad
        entry<float, 1024, 4, BLUE, false, false, false, radices<9, 5, 7, 8, 8>, radices<8, 9, 5, 7, 8>>(),
        // the convolution over a pair of workgroups per transform, for batches too small to fill
        // the device: half-length FFTs of 10240 = 2^11 5 (n <= 10240).  The planner takes it
        // over the smallest 7-smooth half length when it is at most 3% longer (n = 10007: 10080,
        // 9 5 7 4 8, five passes with ragged rounds): four passes whose butterfly rounds fill
        // the 1024 work-items (2048, 640, 640 and 1280 butterflies), measured 35.0 -> 31.7 us at
        // batch 128; the two LDS buffers (160 KiB, one workgroup per CU either way) save one
        // barrier per exchange, 31.7 -> 29.7 us.  Measured against it at batch 128 (event-timed back
        // to back): twiddle prefetch one pass ahead 35.5-36.3 us (spills), the radix order
        // 5 8 16 16 / 16 16 8 5 30.5-31.1 us; the pad layout needs more than 160 KiB here
        entry<float, 1024, 4, BLUE | SFH_M_PAIR | SFH_M_DBUF, false, false, false, radices<5, 16, 16, 8>,
              radices<8, 16, 16, 5>>(),
        // complex<double> 2520 = 2^3 3^2 5 7: 20 KiB split tiles, 512 work-items (60 VGPRs), four
        // workgroups per CU: 32 waves per CU, the occupancy limit.  At 2520 x 1024 (one round of
        // tiles), interleaved over two rounds: 20.4-20.5 us back to back and 32.4-32.6 us after a
        // cache scrub, against 22.8-23.1 / 36.6-36.9 us for 384 work-items (78 VGPRs), 21.7-22.0 /
        // 32.8-33.1 us for 256; the pad layout and the radix order 5 7 9 8 made no difference
        entry<double, 512, 8, 0, false, true, false, radices<9, 5, 7, 8>, none>(),
        // complex<double> 6561 = 3^8: one radix-9 butterfly per work-item and pass; 52.5 KiB
        // split tiles, two workgroups per CU.  At 6561 x 512: 26.8-27.1 us back to back against
        // 27.3-27.5 us for 1024 work-items and 33.9 us for 832
        entry<double, 768, 6, 0, false, true, false, radices<9, 9, 9, 9>, none>(),
        // complex<float> powers of two, pad layout where linear.  Measured at 512 MiB batches
        // against 2-4x the work-items per transform: fewer work-items with more points each
        // (more independent butterflies per wave between barriers) ran 2-5% faster at every
        // length; the register file then bounds the tiles per CU (about 256 KiB of points).
        // 4096 for batches of at most one transform per CU (latency-bound single round): 256
        // work-items x 16 points, radix 16 16 16 (two exchanges), two LDS buffers (one barrier
        // per exchange).  At 4096 x 256, interleaved over two rounds: 10.8-11.0 us back to back
        // and 14.9 us after a cache scrub, against 11.4-11.7 / 15.4-15.6 us for 512 x 8 points
        // (4 8 16 8) and 11.0-11.2 / 15.4 us for 256 x 16 (32 16 8)
        entry<float, 256, 2, SFH_M_DBUF, true, false, false, radices<16, 16, 16>, none>(1),
        // 4096, larger batches: 128 work-items x 32 points, split tiles of 16.5 KiB.  Radix 32
        // first (32 16 8: two LDS exchanges, every boundary pad-linear) measured 203-214 us per
        // 512 MiB against 215-218 us for 4 8 16 8 (three exchanges)
        entry<float, 128, 4, 0, true, true, false, radices<32, 16, 8>, none>(),
        // 8192: 256 work-items x 32 points, 32 16 16 (two exchanges, pad-linear): 209-210 us
        // against 214 us for 16 2 16 16
        entry<float, 256, 2, 0, true, true, false, radices<32, 16, 16>, none>(),
        // 16384: 512 work-items x 32 points (two tiles per CU).  Measured at 512 MiB against
        // 256 x 64 (16 16 16 4, 223 VGPRs): 0.54 against 0.50-0.52 of the HBM peak; the radix
        // orders 8 16 16 8 and 16 8 16 8 ran within 1%, twiddle prefetch spilled (0.44); with
        // radix 32, 32 32 16 12% slower.  With the factored pass twiddles 32 16 32 (two
        // exchanges; 8 VGPRs spilled) runs 225.0-230.3 us against 228.6 us for 16 8 8 16 (three
        // exchanges, next entry) and 226.9-229.1 us for 16 32 32
        entry<float, 512, 4, 0, true, true, false, radices<32, 16, 32>, none>(),
        entry<float, 512, 4, 0, true, true, false, radices<16, 8, 8, 16>, none>(),
        // 32768 in one pass: with the split exchange one plane (128 KiB + pad) fits the LDS, and
        // the 64 points of each work-item stay in registers across every exchange; 32 32 32 (two
        // exchanges) measured 302 us against 306-307 us for 16 16 16 8, and 1024 x 32 points
        // 303-307 us (with the factored pass twiddles: 279.5-286.5 against 280.2-281.5 us)
        entry<float, 512, 2, 0, true, true, false, radices<32, 32, 32>, none>(),
        // complex<double> 4096 and 2048: 16 points per work-item, split tiles (33 and 16.5 KiB).
        // At 512 MiB batches: 4096 196.5-196.9 us (0.68 of the HBM peak) against 218-219 us on
        // the generic kernel; 2048 195.6 us against 209.3-209.6 us.  (fp32 2048 on 128 x 16
        // points or 64 x 32 points tied the generic kernel, 199-222 against 201-217 us.)
        entry<double, 256, 4, 0, true, true, false, radices<16, 16, 16>, none>(),
        entry<double, 128, 4, 0, true, true, false, radices<16, 16, 8>, none>(),
        // complex<float> 2048 (tied the generic kernel as a plain entry; present for its C2R form)
        entry<float, 64, 4, 0, true, true, false, radices<32, 16, 4>, none>(),
        // complex<float> 1024: one wave x 16 points, for real plans of N = 2048 only (at 512 MiB
        // of complex values 210-218 us against 204-215 us on the generic kernel; its C2R / R2C
        // forms remove a pass)
        entry<float, 64, 4, 0, true, true, false, radices<16, 16, 4>, none>(0, 1),
        // C2R real plans: the tangle of the half spectrum fused into the first pass of the
        // length-M backward FFT (one HBM pass instead of two).  Each point is loaded once and
        // its mirror exchanged between lanes l and l ^ 32: against loading both X[j] and
        // X[L - j] per point (1.5x the input read from HBM) f64 4096 x 8192 0.134 -> 0.121-0.123
        // ms, real 2-D 2048^2 x 16 0.334 -> 0.326 ms, fp32 4096 x 16384 unchanged (0.148-0.151).
        // Same radix programs as the plain entries of these lengths, whose pass tables they share
        entry<double, 128, 4, SFH_M_C2R, true, true, false, radices<16, 16, 8>, none>(),
        entry<double, 256, 4, SFH_M_C2R, true, true, false, radices<16, 16, 16>, none>(),
        entry<float, 64, 4, SFH_M_C2R, true, true, false, radices<32, 16, 4>, none>(),
        entry<float, 128, 4, SFH_M_C2R, true, true, false, radices<32, 16, 8>, none>(),
        entry<float, 256, 2, SFH_M_C2R, true, true, false, radices<32, 16, 16>, none>(),
        entry<float, 64, 4, SFH_M_C2R, true, true, false, radices<16, 16, 4>, none>(),
        // R2C real plans: the untangle fused into the last pass (one more split LDS exchange),
        // same radix programs as the plain entries.  The complex<double> forms at three waves
        // per SIMD (132 VGPRs, no spills): f64 4096 x 8192 0.1116 ms against 0.113-0.114 ms at
        // four (4-6 VGPRs spilled, 6% more HBM traffic)
        entry<double, 128, 3, SFH_M_R2C, true, true, false, radices<16, 16, 8>, none>(),
        entry<double, 256, 3, SFH_M_R2C, true, true, false, radices<16, 16, 16>, none>(),
        entry<float, 64, 4, SFH_M_R2C, true, true, false, radices<32, 16, 4>, none>(),
        entry<float, 128, 4, SFH_M_R2C, true, true, false, radices<32, 16, 8>, none>(),
        entry<float, 256, 2, SFH_M_R2C, true, true, false, radices<32, 16, 16>, none>(),
        entry<float, 64, 4, SFH_M_R2C, true, true, false, radices<16, 16, 4>, none>(),
        // column tiles (strided axes), with the group of adjacent tiles run on one XCD (rows
        // of 64 bytes share cache lines).  Groups of 1 (no grouping), 2, 4, 8 and 16 tiles,
        // interleaved over two rounds: C3 complex<double> 4096^2 in place 228-257 / 199-229 /
        // 198-199 / 216-219 / 203-205 us; complex<float> 2048^2 x 16 in place 562-564 / 543-546
        // / 541 / 517-518 / 521 us; real 2048^2 x 16 R2C 356-364 / 325-328 / 311-313 / 307-311
        // / 307 us, C2R 409-411 / 376-377 / 351-352 / 341 / 333-334 us (1025 half-spectrum
        // columns: unaligned rows)
        // complex<double> 4096 x 4 columns, 1024 work-items x 16 points (256 per column), split
        // regions of 33 KiB per column, groups of 4 tiles
        entry<double, 1024, 4, 0, true, true, false, radices<16, 16, 16>, none, 4>(0, 0, 2),
        // complex<float> 2048 x 8 columns, 1024 work-items x 16 points, split regions of 8.3 KiB
        // per column, groups of 8 tiles.  16 columns per tile (rows of 128 bytes, 1024 x 32
        // points, one tile per CU, groups of 4) measured slower: real 2-D 2048^2 x 16 R2C
        // 316-323 against 291-293 us, C2R 351-355 against 309-315 us, fp32 2048^2 x 16 526
        // against 517-518 us
        entry<float, 1024, 4, 0, true, true, false, radices<16, 16, 8>, none, 8>(0, 0, 3),
    };
    *count = static_cast<int>(sizeof(k) / sizeof(k[0]));
    return k;
}
} // namespace

extern "C" const sfh_fixed_kernel *sfh_fixed_kernels(int *count)
{
    return table(count);
}
