// Kernel-argument layout shared by the plan compiler (host) and the tile kernels (device).
//
// One launch of the tile kernel performs, for every "column" of a 2-D tile held in LDS,
// one FFT of length L (optionally followed by a pointwise product and a second FFT of
// the same length), with fused pre-ops on load and post-ops on store.  Every transform
// the reference executes with one or more `parallel_for` passes
// (src/sycl/sycl.cpp:87-339: copy_scaled, mixed-radix axis, radix-2 axis, Bluestein axis)
// is expressed as one to three launches of this kernel; the plan compiler in
// fft_plan.cpp chooses the tiles, strides and fusions.
#pragma once

#include <stdint.h>

#define SFH_MAX_PASSES 32 /* FFT #1 passes followed by FFT #2 passes */

// Division by a runtime constant 1 <= d < 2^31 for dividends n < 2^31:
// q = (umulhi(n, m) + n) >> s   with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
struct sfh_fastdiv
{
    uint32_t d;
    uint32_t m;
    uint32_t s;
};

// One Stockham radix-r pass.  `ns` is the product of the radices already applied;
// `twstride` = L / (ns * r) maps the pass twiddle exponent j*t onto the length-L table.
struct sfh_pass
{
    uint32_t radix;
    uint32_t ns;
    uint32_t twstride;
    sfh_fastdiv ns_div;
    // LDS layout of the tile this pass reads (pad_in) and writes (pad_out): 0 = element e
    // at slot e, 1 = one pad slot after every 32 elements, e at e + e/32
    uint32_t pad_in;
    uint32_t pad_out;
};

// Compile-time fusion modes (template parameter of the kernel).
enum sfh_mode : int
{
    SFH_M_PRE = 1,  // multiply loaded element j by pre[p], zero-fill p >= nvalid_in
    SFH_M_MID = 2,  // after FFT #1: multiply by mid[s], conjugate, run FFT #2
    SFH_M_T4 = 4,   // after the last FFT: multiply output k by W_N^(col * k)
    SFH_M_POST = 8, // multiply stored element by post[q], skip q >= nvalid_out
    // fp32 plans: radix-16 passes as 16-point DFT matrix products on MFMA with split
    // operands (hi + lo, three products, FP32 accumulation) instead of VALU butterflies
    SFH_M_SPLIT_BF16 = 16, // BF16 hi/lo
    SFH_M_SPLIT_F16 = 32,  // FP16 hi/lo, each butterfly scaled by a power of two
    // length-specialized Bluestein over two workgroups per transform (with PRE | MID | POST):
    // each writes its half's partial result to pair_buf, and the second of the pair to finish
    // adds the two and writes the output
    SFH_M_PAIR = 64,
    // length-specialized kernels: two LDS buffers used in turn by successive exchanges, so an
    // exchange needs one barrier (between its writes and reads) instead of two
    SFH_M_DBUF = 128,
    // length-specialized kernels, C2R tangle on load (real plans): point j of the length-L
    // complex FFT is Z[j] = (X[j] + conj X[L-j]) + i W_2L^-j (X[j] - conj X[L-j]) of the
    // L + 1-point half spectrum X at the tile's source, with pre[j] = W_2L^j
    SFH_M_C2R = 256,
    // length-specialized kernels, R2C untangle on store (real plans): the length-L FFT Z of
    // the real input read as L complex values goes through LDS, and output k = 0..L of the
    // tile's destination is X[k] = 1/2 (Z[k] + conj Z[L-k]) - i/2 W_2L^k (Z[k] - conj Z[L-k])
    // (Z[L] = Z[0]), with pre[k] = W_2L^k
    SFH_M_R2C = 512
};

// Run-time flags.
enum sfh_flags : uint32_t
{
    SFH_CONJ_IN = 1u << 0,   // conjugate every loaded element (backward transforms)
    SFH_WFAST = 1u << 4,     // lane mapping: consecutive lanes take consecutive columns
    SFH_LOAD_LDS = 1u << 5,  // load through a coalescing LDS loop instead of into pass 0
    SFH_STORE_LDS = 1u << 6, // store through a coalescing LDS loop instead of from the last pass
    SFH_TILE_REVERSE = 1u << 7, // workgroup g runs tile ntiles - 1 - g (last tiles first)
    // column-tile length-specialized launches, bits 8-10: lg = log2 of the group of adjacent
    // tiles (whose rows share cache lines) run on one XCD, as workgroups g, g + 8, g + 16, ...
    // (workgroups are dealt to the 8 XCDs in turn); 0: no grouping, 1: pairs
    SFH_TILE_XCD_SHIFT = 8,
    SFH_TILE_XCD_MASK = 7u << 8,
    SFH_TILE_XCD_PAIR = 1u << 8
};

// Dimension of a launch's tile grid that enumerates transforms (for unit tables).
enum sfh_unit_dim : uint32_t
{
    SFH_UNIT_COL = 1, // column index w0 + w (contiguous 1-D axes: one column per line)
    SFH_UNIT_BLO = 2, // low batch index blo
    SFH_UNIT_BHI = 3  // high batch index bhi
};

// Load pipeline:  read -> CONJ_IN -> [PRE: zero if p >= nvalid_in, else * pre[p]]
// Mid pipeline:   * mid[s] -> conjugate -> FFT #2
// Store pipeline: [T4] -> [POST: skip if q >= nvalid_out, else * post[q]]
//                 -> (re * out_sx, im * out_sy) -> write
// The planner folds conjugations and the normalization scale into (out_sx, out_sy): a
// conjugation is out_sy = -out_sx; a conjugation before POST (the Bluestein inverse FFT's
// conj-FFT-conj) is folded in with a conjugated post table, conj(v) p = conj(v conj(p)).

// with p = j * pre_jmul + col * pre_cmul, s = k * mid_kmul + col * mid_cmul,
//      q = k * post_kmul + col * post_cmul, T4 exponent = col * k,
//      col = (wg / col_div) % col_mod + blo * col_bmul, wg = global column index.

struct sfh_tile_args
{
    const void *in;
    void *out;
    // Twiddle tables staged into LDS at kernel start (tw_entries + t4_entries values):
    //   pass twiddles W_L^m = hi[m >> tw_bits] * lo[m & (2^tw_bits - 1)]
    //     lo[x] = W_L^x (x < 2^tw_bits), hi[y] = W_L^(y 2^tw_bits)
    //   four-step twiddles W_N^m = T0[d0] * T1[d1] * T2[d2], base-2^t4_bits digits of m
    //     T0[x] = W_N^x, T1[x] = W_N^(x 2^b), T2[y] = W_N^(y 2^2b)
    const void *twtab;
    const void *pre;  // pre-op table
    const void *mid;  // mid-op table (spectrum with 1/M folded in)
    const void *post; // post-op table

    // Per-transform unit tables (execution over pointer arrays): when set, the base of unit
    // u of the launch's transform dimension (unit_dim) on that side is units[u] instead of
    // in/out + u * stride.  Null for plan-owned scratch and for ordinary executions.
    const void *const *in_units;
    void *const *out_units;

    // tile index g -> (wt, blo, bhi): wt fastest, then blo, then bhi
    uint64_t wcount; // extent of the column dimension (ragged last tile is masked)
    uint64_t nwt;    // number of column tiles = ceil(wcount / W)
    uint64_t nblo;   // extent of the low batch dimension
    uint64_t ntiles; // nwt * nblo * nbhi
    int64_t in_blo, in_bhi, out_blo, out_bhi; // element strides of the batch dimensions
    int32_t in_w, in_j, out_w, out_k;         // element strides inside a tile (< 2^31 span)

    uint32_t L;      // FFT length held in LDS
    uint32_t LP;     // padded LDS column pitch (elements)
    uint32_t W;      // columns per tile (power of two)
    uint32_t logW;
    uint32_t tpc;    // threads per column = NT / W
    uint32_t logtpc;
    uint32_t npass;  // passes of FFT #1 (FFT #2 uses the reversed radix order)
    // passes per radix in the fixed kernel order 16, 8, 4, 2, 9, 3, 5, 7, without pass 0
    // (FFT #1 runs them in this order after pass 0); rcount2: FFT #2's passes without its
    // first one, which is fused into FFT #1's last (FFT #2 runs them in the reverse order)
    uint8_t rcount[8];
    uint8_t rcount2[8];
    uint32_t flags;
    uint32_t unit_dim; // SFH_UNIT_*: the dimension that enumerates transforms
    sfh_fastdiv L_div;
    sfh_fastdiv col_div, col_mod;
    uint32_t col_bmul;
    uint32_t pre_jmul, pre_cmul, nvalid_in;
    uint32_t mid_kmul, mid_cmul;
    uint32_t post_kmul, post_cmul, nvalid_out;
    uint32_t tw_bits, tw_entries; // pass-twiddle tables (lo has 2^tw_bits entries)
    uint32_t t4_bits, t4_entries; // four-step tables (T0, T1 have 2^t4_bits entries)
    uint32_t lds_tab;             // LDS element offset of the tables (after the W x LP tile)
    uint32_t pad_store;           // LDS layout of the tile the STORE_LDS loop reads
    uint32_t lds_bytes;           // dynamic LDS of the launch
    double out_sx, out_sy; // store epilogue: re * out_sx, im * out_sy
    // SFH_M_PAIR: partial results, slot (2 t + h) of pair_stride points for half h of transform
    // t, and one arrival counter per transform (even before an execution: each adds 2)
    void *pair_buf;
    uint32_t *pair_count;
    uint32_t pair_stride;
    // The plan's table block: the length-specialized kernels read their tables (pass
    // twiddles, chirps, spectra) from global memory, pass after pass, and touch this block
    // at their start so that it is in L2 by the time later passes need it (null: no touch).
    const void *warm;
    uint64_t warm_bytes;
    sfh_pass pass[SFH_MAX_PASSES];
};

// Work-queue launches: one persistent launch runs the tiles of several single-launch plans
// of one kernel variant (plain fusion mode).  Item k is tile `tile` of the launch whose
// arguments are args[arg]; each workgroup takes the next item from *counter until none is
// left, so plans of different lengths, batch counts and directions share one dispatch.
struct sfh_queue_item
{
    uint32_t arg;
    uint32_t tile;
};

struct sfh_queue_args
{
    const struct sfh_tile_args *args;
    const struct sfh_queue_item *items;
    uint32_t *counter; // 0 when the launch starts (written by the upload before it)
    uint32_t nitems;
    uint32_t lds_next; // byte offset in dynamic LDS of the item-index slot (after the largest tile)
};

