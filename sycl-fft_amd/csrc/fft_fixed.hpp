// Length-specialized tile kernels: one transform of a fixed length L per workgroup, with
// the whole radix program known at compile time.
//
// The generic tile kernel (fft_tile.hpp) reads its radix program, strides and twiddle
// layout from the launch arguments, so every pass pays for run-time index arithmetic
// (divisions by ns, two-level twiddle lookups in LDS and the products that rebuild
// W^(j t)).  Here the lengths, radices, ns and butterfly counts are template constants:
//   * butterfly indices, Stockham write positions and LDS offsets fold to constants
//     (a pass costs one base address per butterfly),
//   * pass twiddles come from a per-pass table in global memory (L2-resident, a few
//     hundred KiB at most), laid out [t - 1][j] so lane j reads W_{ns R}^(j t) with one
//     coalesced load per t: no VALU work and no LDS traffic for twiddles,
//   * the first pass reads straight from HBM and the last writes straight to HBM, as in
//     the generic kernel, with the same fused pre / mid / post operations, so a
//     specialized launch is a drop-in for a generic launch of the same arguments.
// Supported: one column per tile (W = 1), contiguous points (in_j = out_k = 1), fusion
// modes plain (0), Bluestein fused (PRE | MID | POST) and Bluestein over a pair of
// workgroups per transform (PRE | MID | POST | PAIR).  The second FFT of the
// Bluestein program has its own radix order (its first radix must be the first FFT's last:
// that pass is fused with the spectrum product in registers).
#pragma once

#include "fft_tile.hpp"

namespace sfh
{
namespace fx
{

template <int... R>
struct radices
{
    static constexpr int n = sizeof...(R);
    static constexpr int v[sizeof...(R)] = {R...};
};

// product of the radices before pass p (the pass's ns)
template <class RL>
constexpr int ns_of(int p)
{
    int s = 1;
    for (int q = 0; q < p; ++q) {
        s *= RL::v[q];
    }
    return s;
}

template <class RL>
constexpr int length_of()
{
    return ns_of<RL>(RL::n);
}

// twiddle entries of passes 1 .. p-1 (pass 0 has ns = 1 and needs none): pass q holds
// ns_q (R_q - 1) values, [t - 1][j] = W_{ns_q R_q}^(j t)
template <class RL>
constexpr int tw_before(int p)
{
    int off = 0;
    for (int q = 1; q < p; ++q) {
        off += ns_of<RL>(q) * (RL::v[q] - 1);
    }
    return off;
}

enum
{
    IN_HBM = 0,
    IN_SHARED = 1
};
enum
{
    OUT_SHARED = 0,
    OUT_HBM = 1,
    OUT_SPECTRUM = 2 // spectrum product, conjugation and the second FFT's first pass, then LDS
};

// One step of the fused program: pass `r` with `ns`, its inputs (HBM or LDS), its outputs
// (LDS, HBM, or the spectrum product + second FFT's first pass), the ns of the pass whose
// outputs it writes (1 after the spectrum product), its twiddle table offset, and the LDS
// layout of its input and output boundaries (pad: point e at slot e + e / 32).
struct step_info
{
    int r, ns, wns, in, out, twoff;
    bool pad_in, pad_out;
};

// Program of the first FFT (radices R1) and, with MID, the second (R2, first pass fused
// into the first FFT's last).  A boundary takes the pad layout when `pad` allows it and
// both sides stay linear in the butterfly's point index t: the writer's points
// base + t ns lie in one 32-point block (ns r divides 32) or ns is a multiple of 32, and
// the reader's points i + t L / r have L / r a multiple of 32 or a divisor of 32.
template <class R1, class R2, int MODE, bool PAD, int W = 1>
struct program
{
    // columns per tile: W = 1 holds one contiguous transform; W > 1 holds W columns of a
    // strided axis, lanes column-fastest (lane l takes column l mod W), each column's points
    // at the launch's point stride, each column in its own LDS region
    static constexpr int cols = W;
    static constexpr int log_cols = W == 1 ? 0 : W == 2 ? 1 : W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : -1;
    static_assert(log_cols >= 0, "columns per tile: 1, 2, 4, 8 or 16");
    static constexpr bool two = (MODE & SFH_M_MID) != 0;
    static constexpr int L = length_of<R1>();
    static constexpr int count = two ? R1::n + R2::n - 1 : R1::n;

    static constexpr step_info raw(int s)
    {
        step_info x{};
        if (s < R1::n) {
            x.r = R1::v[s];
            x.ns = ns_of<R1>(s);
            x.twoff = tw_before<R1>(s);
            x.in = s == 0 ? IN_HBM : IN_SHARED;
            x.out = s + 1 < R1::n ? OUT_SHARED : (two ? OUT_SPECTRUM : OUT_HBM);
        } else {
            const int p = s - R1::n + 1;
            x.r = R2::v[p];
            x.ns = ns_of<R2>(p);
            x.twoff = tw_before<R1>(R1::n) + tw_before<R2>(p);
            x.in = IN_SHARED;
            x.out = s + 1 < count ? OUT_SHARED : OUT_HBM;
        }
        x.wns = x.out == OUT_SPECTRUM ? 1 : x.ns;
        return x;
    }
    static constexpr bool writer_linear(const step_info &w)
    {
        return w.wns % 32 == 0 || (w.wns * w.r <= 32 && 32 % (w.wns * w.r) == 0);
    }
    static constexpr bool reader_linear(const step_info &r)
    {
        const int m = L / r.r;
        return m % 32 == 0 || (m <= 32 && 32 % m == 0);
    }
    // boundary before step s (1 <= s < count)
    static constexpr bool padded(int s)
    {
        return PAD && s >= 1 && s < count && writer_linear(raw(s - 1)) && reader_linear(raw(s));
    }
    static constexpr step_info step(int s)
    {
        step_info x = raw(s);
        x.pad_in = padded(s);
        x.pad_out = padded(s + 1);
        return x;
    }
    // Steps whose butterflies come in pairs i, M - i held by lanes l and l ^ 32 of one wave
    // (the tangle partners: point L - (i + t M) is point R - 1 - t of butterfly M - i): the
    // first step of C2R programs.  (The R2C untangle after the last step goes through LDS:
    // the same pairing there, with the mirror outputs exchanged between lanes, measured
    // 0.155-0.159 / 0.135-0.137 ms against 0.147-0.151 / 0.123-0.124 ms for fp32 4096 x 16384
    // / f64 4096 x 8192 through LDS with whole-value exchanges.)
    static constexpr bool paired(int s)
    {
        return (MODE & SFH_M_C2R) != 0 && s == 0;
    }
    // LDS points (slots) the tile needs
    static constexpr int slots()
    {
        bool any = false;
        for (int s = 1; s < count; ++s) {
            any = any || padded(s);
        }
        return any ? L + (L - 1) / 32 + 1 : L;
    }
    // LDS slots between the regions of two columns (elements of `bytes` each): at least
    // slots(), and W regions start 64 / W banks (4-byte words, 64 banks) apart, so that the
    // lanes of one row (consecutive columns) meet different banks
    static constexpr int column_pitch(int bytes)
    {
        if (W == 1) {
            return slots();
        }
        int cp = slots();
        while ((cp * (bytes / 4)) % 64 != 64 / W) {
            ++cp;
        }
        return cp;
    }
};

// The work-item's lane inside its column (the whole workgroup for W = 1).
template <class PROG>
__device__ __forceinline__ uint32_t lane()
{
    return static_cast<uint32_t>(threadIdx.x) >> PROG::log_cols;
}

// slot offset of point base + off inside a pad-linear run: off + off / 32
template <bool PAD>
constexpr uint32_t run_offset(int off)
{
    return static_cast<uint32_t>(PAD ? off + off / 32 : off);
}
template <bool PAD>
__device__ __forceinline__ uint32_t base_slot(uint32_t e)
{
    return PAD ? e + (e >> 5) : e;
}

// Caller buffers are global memory: say so, so the compiler emits global (not flat)
// accesses, whose completion counts do not mix with LDS operations.
#if defined(__HIP_DEVICE_COMPILE__)
#define SFH_GLOBAL_AS __attribute__((address_space(1)))
#else
#define SFH_GLOBAL_AS
#endif
#if defined(__HIP__)
// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt at their maxima): every vector
// memory access of the wave, stores included, has completed
#define SFH_WAIT_VMEM() __builtin_amdgcn_s_waitcnt(0x0F70)
#else
#define SFH_WAIT_VMEM() __atomic_thread_fence(__ATOMIC_SEQ_CST)
#endif
template <typename T>
__device__ __forceinline__ const SFH_GLOBAL_AS cx<T> *in_global(const cx<T> *p)
{
    return (const SFH_GLOBAL_AS cx<T> *)p;
}
template <typename T>
__device__ __forceinline__ SFH_GLOBAL_AS cx<T> *out_global(cx<T> *p)
{
    return (SFH_GLOBAL_AS cx<T> *)p;
}

template <typename T, int MODE, int COLS = 1>
__device__ __forceinline__ cx<T> load_point(const sfh_tile_args &a, const tile_ctx<T> &c, uint32_t j)
{
    const uint32_t at = COLS > 1 ? j * static_cast<uint32_t>(a.in_j) : j;
    if constexpr (COLS > 1) {
        if (!c.wvalid) { // a column past the last of a ragged tile: no access
            return {T(0), T(0)};
        }
    }
    if constexpr (MODE & SFH_M_PRE) {
        const uint32_t p = j * a.pre_jmul + c.col * a.pre_cmul;
        if (p >= a.nvalid_in) {
            return {T(0), T(0)};
        }
        cx<T> v = in_global(c.src)[at];
        if (a.flags & SFH_CONJ_IN) {
            v.y = -v.y;
        }
        return cmul(v, static_cast<const cx<T> *>(a.pre)[p]);
    } else {
        cx<T> v = in_global(c.src)[at];
        if (a.flags & SFH_CONJ_IN) {
            v.y = -v.y;
        }
        return v;
    }
}

// Butterfly counts of step S of program PROG: M = L / R butterflies, NB per work-item.
template <int NT, int L, class PROG, int S>
struct shape
{
    static constexpr step_info st = PROG::step(S);
    static constexpr int R = st.r;
    static constexpr int M = L / R;
    static constexpr int NB = (M + NT - 1) / NT;
    // butterfly q of a work-item exists (compile-time true except in a partial last round)
    static __device__ __forceinline__ bool has(int q, uint32_t i)
    {
        return (q + 1) * NT <= M || i < static_cast<uint32_t>(M);
    }
};

// Butterfly q of work-item tid in step S: tid + q NT, or in a paired step (NB NT = M, whole
// waves) pair x = q NT / 2 + 32 (tid / 64) + (tid mod 32), whose low member x goes to lanes
// 0-31 and high member M - x to lanes 32-63 of the wave; pair 0 is {0, M / 2} (both their
// own mirrors).
template <int NT, int L, class PROG, int S>
__device__ __forceinline__ uint32_t bfly(uint32_t tid, int q)
{
    using sh = shape<NT, L, PROG, S>;
    if constexpr (PROG::paired(S)) {
        static_assert(NT % 64 == 0 && sh::NB * NT == sh::M && sh::M % 2 == 0,
                      "paired steps: whole waves, every work-item NB butterflies");
        const uint32_t l = tid & 63u;
        const uint32_t x = static_cast<uint32_t>(q * (NT / 2)) + ((tid >> 6) << 5) + (l & 31u);
        return l < 32u ? x : (x == 0u ? static_cast<uint32_t>(sh::M / 2) : static_cast<uint32_t>(sh::M) - x);
    } else {
        return tid + static_cast<uint32_t>(q * NT);
    }
}

// The value of lane l ^ 32 of the wave (every lane of the wave takes part).
template <typename T>
__device__ __forceinline__ cx<T> partner32(const cx<T> &v)
{
    return {wave_xor(v.x, 32), wave_xor(v.y, 32)};
}

template <typename T, int NT, int L, class PROG, int S>
struct regs
{
    using sh = shape<NT, L, PROG, S>;
    cx<T> v[sh::NB][sh::R];
};

// Pass twiddles of step S for the work-item's butterflies, w[q][t - 1] = W_{NS R}^((i mod NS) t).
// They are loaded one step ahead (before the previous step's butterflies and LDS exchange),
// so their L2 latency overlaps that work instead of stalling every wave after the barrier.
template <typename T, int NT, int L, class PROG, int S>
struct twregs
{
    using sh = shape<NT, L, PROG, S>;
    static constexpr bool used = sh::st.in == IN_SHARED && sh::st.ns > 1;
    cx<T> w[used ? sh::NB : 1][used ? sh::R - 1 : 1];
};

template <typename T, int NT, int L, class PROG, int S>
__device__ __forceinline__ void load_twiddles(const cx<T> *tw0, twregs<T, NT, L, PROG, S> &w)
{
    using sh = shape<NT, L, PROG, S>;
    if constexpr (twregs<T, NT, L, PROG, S>::used) {
        constexpr int R = sh::R, NS = sh::st.ns;
        const cx<T> *tw = tw0 + sh::st.twoff;
        const uint32_t tid = lane<PROG>();
#pragma unroll
        for (int q = 0; q < sh::NB; ++q) {
            const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
            if (sh::has(q, i)) {
                const uint32_t j = i % static_cast<uint32_t>(NS);
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    w.w[q][t - 1] = tw[(t - 1) * NS + j];
                }
            }
        }
    }
}

// Step S of a Stockham program on the work-item's butterflies i = tid + q NT (i < L/R):
// inputs t of butterfly i are points i + t L/R; input t is multiplied by W_{NS R}^((i mod NS) t)
// (steps that read LDS), the radix-R DFT runs, and output t goes to point
// (i / NS) NS R + (i mod NS) + t NS of the next step.
// C2R tangle of one point: Z[j] = (X[j] + conj X[L-j]) + i W^-j (X[j] - conj X[L-j]),
// W = W_2L, pre[j] = W^j, conjugated for a backward plan.
template <typename T>
__device__ __forceinline__ cx<T> tangle(const sfh_tile_args &a, const cx<T> &x, cx<T> m, cx<T> w)
{
    m.y = -m.y;
    const cx<T> s = x + m, d = x - m;
    w.y = -w.y;
    const cx<T> t = cmul(w, d);
    cx<T> z = {s.x - t.y, s.y + t.x};
    if (a.flags & SFH_CONJ_IN) {
        z.y = -z.y;
    }
    return z;
}

// C2R tangle in a paired first step: butterfly i loads X[j], j = i + t M, once.  The mirror
// X[L - j] is point R - 1 - t of the partner lane's butterfly M - i (of its own butterfly
// for i = M / 2), so points t and R - 1 - t are tangled together in place.  Butterfly 0 (lane
// 0 of the first wave) mirrors point t in its own point R - t and point 0 in X[L]: it keeps
// its loads through the common loop and is tangled after it.
template <typename T, int L, int R, int M>
__device__ __forceinline__ void load_c2r_paired(const sfh_tile_args &a, const tile_ctx<T> &c, uint32_t i, cx<T> *v)
{
    const SFH_GLOBAL_AS cx<T> *src = in_global(c.src);
#pragma unroll
    for (int t = 0; t < R; ++t) {
        v[t] = src[i + static_cast<uint32_t>(t * M)];
    }
    cx<T> xl = {T(0), T(0)};
    if (i == 0u) {
        xl = src[L];
    }
    // W^(i + t M) = W^i W^(t M): one table load per butterfly, the W^(t M) the same for every
    // lane (scalar loads)
    const cx<T> *pre = static_cast<const cx<T> *>(a.pre);
    const cx<T> wi = pre[i];
    const bool half = i == static_cast<uint32_t>(M / 2);
#pragma unroll
    for (int p = 0; p < R / 2; ++p) {
        const int u = R - 1 - p;
        const cx<T> x0 = v[p], x1 = v[u];
        cx<T> m0 = partner32(x1), m1 = partner32(x0);
        if (half) {
            m0 = x1;
            m1 = x0;
        }
        const cx<T> z0 = tangle(a, x0, m0, cmul(wi, pre[p * M]));
        const cx<T> z1 = tangle(a, x1, m1, cmul(wi, pre[u * M]));
        if (i != 0u) {
            v[p] = z0;
            v[u] = z1;
        }
    }
    if (i == 0u) {
        const cx<T> z0 = tangle(a, v[0], xl, pre[0]);
        const cx<T> zh = tangle(a, v[R / 2], v[R / 2], pre[(R / 2) * M]);
#pragma unroll
        for (int t = 1; t < R / 2; ++t) {
            const cx<T> x0 = v[t], x1 = v[R - t];
            v[t] = tangle(a, x0, x1, pre[t * M]);
            v[R - t] = tangle(a, x1, x0, pre[(R - t) * M]);
        }
        v[0] = z0;
        v[R / 2] = zh;
    }
}

template <bool PF, typename T, int NT, int L, int MODE, class PROG, int S>
__device__ __forceinline__ void compute(const sfh_tile_args &a, const tile_ctx<T> &c, const cx<T> *tw0,
                                        const twregs<T, NT, L, PROG, S> &w, regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr step_info st = sh::st;
    constexpr int R = sh::R, M = sh::M, NS = st.ns;
    const uint32_t tid = lane<PROG>();
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
        if (sh::has(q, i)) {
            cx<T> *v = r.v[q];
            if constexpr (st.in == IN_HBM && PROG::paired(S)) {
                load_c2r_paired<T, L, R, M>(a, c, i, v);
            } else if constexpr (st.in == IN_HBM) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[t] = load_point<T, MODE, PROG::cols>(a, c, i + static_cast<uint32_t>(t * M));
                }
            } else if constexpr (NS > 1 && !PF && (sizeof(T) == 4 || R % 2 == 1) && R >= 5) {
                // complex<float> radix >= 5, complex<double> odd radix >= 5: load W^(j 2^b)
                // (rows 2^b - 1 of the pass table)
                // and form W^(j t) = W^(j h) W^(j (t - h)), h the highest power of two in t:
                // floor(log2 (R - 1)) + 1 loads per butterfly instead of R - 1, at most
                // log2 R - 1 products deep.  At 512 MiB batches, against one load per twiddle
                // (power-of-two radices): 4096 206-209 against 217-219 us, 8192 205 / 216-219,
                // 16384 229-231 / 244-245, 32768 280-282 / 296-297, the Bluestein pair (10007 x
                // 128) 28.2-28.3 / 29.7-29.9 us; its radix-5 passes too: 28.3-28.4 / 28.8-28.9 us.
                // complex<double> takes it for the odd radices (6561 x 512 26.3-26.7 against
                // 27.1-27.4 us, 2520 x 1024 20.0-20.1 against 20.3-20.9 us); with radix 16 the
                // products made it slower (4096: 210 against 196 us), so its power-of-two passes
                // keep one table load per twiddle
                const uint32_t j = i % static_cast<uint32_t>(NS);
                const cx<T> *tw = tw0 + st.twoff;
                cx<T> wp[R];
#pragma unroll
                for (int h = 1; h < R; h *= 2) {
                    wp[h] = tw[(h - 1) * NS + j];
                }
#pragma unroll
                for (int t = 3; t < R; ++t) {
                    if ((t & (t - 1)) != 0) {
                        int h = 1;
                        while (h * 2 <= t) {
                            h *= 2;
                        }
                        wp[t] = cmul(wp[h], wp[t - h]);
                    }
                }
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    v[t] = cmul(v[t], wp[t]);
                }
            } else if constexpr (NS > 1) {
                // prefetched (PF) or loaded here, per butterfly (fewer live registers)
                const uint32_t j = i % static_cast<uint32_t>(NS);
                const cx<T> *tw = tw0 + st.twoff;
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    v[t] = cmul(v[t], PF ? w.w[q][t - 1] : tw[(t - 1) * NS + j]);
                }
            }
            dft<T, R>::run(v);
            if constexpr (st.out == OUT_SPECTRUM) {
                // natural-order outputs k = i + t M of the first FFT: times the spectrum,
                // conjugated (the inverse FFT as conj(FFT(conj))), then the second FFT's first
                // pass (radix R, ns = 1) on the same points
                const cx<T> *mid = static_cast<const cx<T> *>(a.mid);
                const uint32_t s0 = i * a.mid_kmul + c.col * a.mid_cmul;
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[t] = cconj(cmul(v[t], mid[s0 + static_cast<uint32_t>(t * M) * a.mid_kmul]));
                }
                dft<T, R>::run(v);
            }
        }
    }
}

// Point access that is coherent across the device without cache maintenance: agent-scope
// relaxed atomic accesses, which gfx950 performs with the sc1 (coherent) cache policy, so a
// partial result written on one XCD is read from memory, not from a stale L2 line, on
// another.  (A device-scope fence would write back and invalidate whole L2 caches instead:
// measured at ~1 us per workgroup.)
template <typename T>
__device__ __forceinline__ void coherent_put(cx<T> *p, const cx<T> &v)
{
    static_assert(sizeof(cx<T>) % 8 == 0, "complex of 32- or 64-bit parts");
    uint64_t w[sizeof(cx<T>) / 8];
    __builtin_memcpy(w, &v, sizeof(w));
#pragma unroll
    for (int k = 0; k < static_cast<int>(sizeof(cx<T>) / 8); ++k) {
        __hip_atomic_store(reinterpret_cast<uint64_t *>(p) + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
template <typename T>
__device__ __forceinline__ cx<T> coherent_get(const cx<T> *p)
{
    uint64_t w[sizeof(cx<T>) / 8];
#pragma unroll
    for (int k = 0; k < static_cast<int>(sizeof(cx<T>) / 8); ++k) {
        w[k] = __hip_atomic_load(reinterpret_cast<const uint64_t *>(p) + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    cx<T> v;
    __builtin_memcpy(&v, w, sizeof(v));
    return v;
}

// The store paths below issue all of a butterfly group's loads (tables, partner partials)
// before its first store: CDNA counts vector loads and stores on one in-order counter, so a
// load placed after a store waits for that store's completion too, and the compiler cannot
// hoist a table load above a store to a buffer that might alias it.  Interleaved, every
// point would cost a full memory round trip.

// Two workgroups per transform (SFH_M_PAIR): half h of transform b holds its partial result
// P_h[k] = v[k] post[2k + h] (k < n) in registers.  Both halves store theirs to slot 2b + h
// of pair_buf with coherent stores, wait until every store of the workgroup has completed
// (each wave waits for its stores, then the barrier), and count their arrival on
// pair_count[b]; the half that arrives second (odd count: each execution adds 2) reads the
// partner's slot with coherent loads, issued only after it has seen the count, and writes
// ep(P_0 + P_1) to the output.  The sum is the same whichever half adds it (floating-point
// addition commutes).  Nothing waits on another workgroup, so no arrival order can stall.
template <typename T, int NT, int L, int MODE, class PROG, int S>
__device__ __forceinline__ void store_pair(const sfh_tile_args &a, const tile_ctx<T> &c, unsigned char *lds,
                                           regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr int R = sh::R, M = sh::M;
    const uint32_t tid = lane<PROG>();
    const uint32_t b = c.w0, h = c.blo;
    cx<T> *const pair = static_cast<cx<T> *>(a.pair_buf);
    cx<T> *const mine = pair + (2 * static_cast<uint64_t>(b) + h) * a.pair_stride;
    const cx<T> *const other = pair + (2 * static_cast<uint64_t>(b) + (h ^ 1u)) * a.pair_stride;
    const cx<T> *post = static_cast<const cx<T> *>(a.post);
    cx<T> pv[sh::NB][R];
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint32_t pq = (i + static_cast<uint32_t>(t * M)) * a.post_kmul + c.col * a.post_cmul;
            pv[q][t] = (sh::has(q, i) && pq < a.nvalid_out) ? post[pq] : cx<T>{T(0), T(0)};
        }
    }
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
        if (sh::has(q, i)) {
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t k = i + static_cast<uint32_t>(t * M);
                const uint32_t pq = k * a.post_kmul + c.col * a.post_cmul;
                if (pq < a.nvalid_out) {
                    r.v[q][t] = cmul(r.v[q][t], pv[q][t]);
                    coherent_put(mine + k, r.v[q][t]);
                }
            }
        }
    }
    SFH_WAIT_VMEM(); // this wave's stores have completed
    __syncthreads(); // ... every wave's (and the last exchange's LDS reads are done)
    uint32_t *const flag = reinterpret_cast<uint32_t *>(lds);
    if (tid == 0) {
        *flag = __hip_atomic_fetch_add(a.pair_count + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u;
    }
    __syncthreads();
    if (*flag == 0u) {
        return; // first of the pair: the partner adds
    }
    const epilogue<T> ep(a);
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
        if (sh::has(q, i)) {
            cx<T> o[R];
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t k = i + static_cast<uint32_t>(t * M);
                const uint32_t pq = k * a.post_kmul + c.col * a.post_cmul;
                o[t] = pq < a.nvalid_out ? coherent_get(other + k) : cx<T>{T(0), T(0)};
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t k = i + static_cast<uint32_t>(t * M);
                const uint32_t pq = k * a.post_kmul + c.col * a.post_cmul;
                if (pq < a.nvalid_out) {
                    out_global(c.dst)[k] = ep(r.v[q][t] + o[t]);
                }
            }
        }
    }
}

// Last step: output t of butterfly i is point i + t M (NS R = L, i < NS): post-op, epilogue, HBM.
template <typename T, int NT, int L, int MODE, class PROG, int S>
__device__ __forceinline__ void store(const sfh_tile_args &a, const tile_ctx<T> &c, const regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr int R = sh::R, M = sh::M;
    const epilogue<T> ep(a);
    const uint32_t tid = lane<PROG>();
    if constexpr (MODE & SFH_M_POST) {
        cx<T> pv[sh::NB][R];
#pragma unroll
        for (int q = 0; q < sh::NB; ++q) {
            const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t pq = (i + static_cast<uint32_t>(t * M)) * a.post_kmul + c.col * a.post_cmul;
                pv[q][t] = (sh::has(q, i) && pq < a.nvalid_out) ? static_cast<const cx<T> *>(a.post)[pq]
                                                                : cx<T>{T(0), T(0)};
            }
        }
#pragma unroll
        for (int q = 0; q < sh::NB; ++q) {
            const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
            if (sh::has(q, i)) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const uint32_t k = i + static_cast<uint32_t>(t * M);
                    if (k * a.post_kmul + c.col * a.post_cmul < a.nvalid_out) {
                        out_global(c.dst)[PROG::cols > 1 ? k * static_cast<uint32_t>(a.out_k) : k] =
                            ep(cmul(r.v[q][t], pv[q][t]));
                    }
                }
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < sh::NB; ++q) {
            const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
            if (sh::has(q, i) && (PROG::cols == 1 || c.wvalid)) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const uint32_t k = i + static_cast<uint32_t>(t * M);
                    out_global(c.dst)[PROG::cols > 1 ? k * static_cast<uint32_t>(a.out_k) : k] = ep(r.v[q][t]);
                }
            }
        }
    }
}

// R2C untangle (SFH_M_R2C): after the last step, work-item j takes the output pairs k, L - k
// for k = j, j + NT, ... <= L / 2, which both come from Z[k] and Z[L - k] (Z[L] = Z[0]):
//   X[k]     = 1/2 (Z[k] + conj Z[L-k]) - i/2 W^k (Z[k] - conj Z[L-k]),  W = W_2L,
//   X[L - k] = the same with Z[k] and Z[L-k] swapped and W^(L-k) = -conj W^k.
// Z goes through the LDS plane of the split exchange in natural order, real parts first
// (each work-item keeps Re Z[k], Re Z[L-k] of its pairs), then imaginary parts; the
// outputs (L + 1 per transform) go through the epilogue (scale).  W^(j + u NT) = W^j
// W^(u NT): one table load per work-item, W^(u NT) uniform over the lanes.
template <typename T, int NT, int L, int MODE, class PROG, int S>
__device__ __forceinline__ void store_r2c(const sfh_tile_args &a, const tile_ctx<T> &c, unsigned char *lds,
                                          const regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr int R = sh::R, M = sh::M;
    constexpr int NP = (L / 2 + NT) / NT; // pairs per work-item (k <= L / 2)
    const uint32_t tid = lane<PROG>();
    T *z = reinterpret_cast<T *>(lds);
    T re[NP][2], im[NP][2];
    __syncthreads(); // the last exchange's LDS reads are done
#pragma unroll
    for (int part = 0; part < 2; ++part) {
#pragma unroll
        for (int q = 0; q < sh::NB; ++q) {
            const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
            if (sh::has(q, i)) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    z[i + static_cast<uint32_t>(t * M)] = part == 0 ? r.v[q][t].x : r.v[q][t].y;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const uint32_t k = tid + static_cast<uint32_t>(u * NT);
            if (k <= static_cast<uint32_t>(L / 2)) {
                T *dst = part == 0 ? re[u] : im[u];
                dst[0] = z[k];
                dst[1] = z[k == 0u ? 0u : static_cast<uint32_t>(L) - k];
            }
        }
        if (part == 0) {
            __syncthreads(); // the real parts are read before the imaginary parts overwrite them
        }
    }
    const epilogue<T> ep(a);
    const cx<T> *w = static_cast<const cx<T> *>(a.pre);
    const cx<T> wj = w[tid];
    const auto untangle = [&](const cx<T> &zk, cx<T> zm, const cx<T> &wk, uint32_t k) {
        zm.y = -zm.y;
        const cx<T> s = zk + zm, d = zk - zm;
        const cx<T> t = cmul(wk, d);
        // 1/2 (s - i t) = 1/2 (s.x + t.y, s.y - t.x)
        out_global(c.dst)[k] = ep(cx<T>{T(0.5) * (s.x + t.y), T(0.5) * (s.y - t.x)});
    };
#pragma unroll
    for (int u = 0; u < NP; ++u) {
        const uint32_t k = tid + static_cast<uint32_t>(u * NT);
        if (k <= static_cast<uint32_t>(L / 2)) {
            const cx<T> zk = {re[u][0], im[u][0]}, zm = {re[u][1], im[u][1]};
            const cx<T> wk = cmul(wj, w[u * NT]);
            untangle(zk, zm, wk, k);
            if (k < static_cast<uint32_t>(L / 2)) {
                untangle(zm, zk, cx<T>{-wk.x, wk.y}, static_cast<uint32_t>(L) - k);
            }
        }
    }
}

// Component access of one point: the whole complex value (PART = 2), or its real (0) or
// imaginary (1) part alone for the split exchange.
template <int PART, typename T, typename U>
__device__ __forceinline__ void put(U *lds, uint32_t slot, const cx<T> &v)
{
    if constexpr (PART == 2) {
        reinterpret_cast<cx<T> *>(lds)[slot] = v;
    } else {
        reinterpret_cast<T *>(lds)[slot] = PART == 0 ? v.x : v.y;
    }
}
template <int PART, typename T, typename U>
__device__ __forceinline__ void get(const U *lds, uint32_t slot, cx<T> &v)
{
    if constexpr (PART == 2) {
        v = reinterpret_cast<const cx<T> *>(lds)[slot];
    } else if constexpr (PART == 0) {
        v.x = reinterpret_cast<const T *>(lds)[slot];
    } else {
        v.y = reinterpret_cast<const T *>(lds)[slot];
    }
}

// Write step S's outputs (component PART) to their LDS slots.
template <int PART, typename T, int NT, int L, class PROG, int S>
__device__ __forceinline__ void write_out(unsigned char *lds, const regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr step_info st = sh::st;
    constexpr int R = sh::R, WNS = st.wns;
    const uint32_t tid = lane<PROG>();
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
        if (sh::has(q, i)) {
            const uint32_t kq = i / static_cast<uint32_t>(WNS);
            const uint32_t b = base_slot<st.pad_out>(kq * static_cast<uint32_t>(WNS * (R - 1)) + i);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                put<PART>(lds, b + run_offset<st.pad_out>(t * WNS), r.v[q][t]);
            }
        }
    }
}

// Read step S's inputs (component PART): points i + t M.
template <int PART, typename T, int NT, int L, class PROG, int S>
__device__ __forceinline__ void read_in(const unsigned char *lds, regs<T, NT, L, PROG, S> &r)
{
    using sh = shape<NT, L, PROG, S>;
    constexpr step_info st = sh::st;
    constexpr int R = sh::R, M = sh::M;
    const uint32_t tid = lane<PROG>();
#pragma unroll
    for (int q = 0; q < sh::NB; ++q) {
        const uint32_t i = bfly<NT, L, PROG, S>(tid, q);
        if (sh::has(q, i)) {
            const uint32_t b = base_slot<st.pad_in>(i);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                get<PART>(lds, b + run_offset<st.pad_in>(t * M), r.v[q][t]);
            }
        }
    }
}

// LDS exchange from step S's outputs to step S + 1's inputs.  Every work-item's values are
// in registers on both sides, so with SPLIT the real parts move first and the imaginary
// parts second through an LDS buffer of one real per point (half the footprint, twice the
// barriers).  The leading barrier ends the previous exchange's reads (none before the first).
// With DBUF, exchange S uses buffer S mod 2: the previous exchange read the other buffer, and
// the exchange before it (same buffer) ended its reads before the previous exchange's
// barrier, so no leading barrier is needed.
template <bool SPLIT, bool DBUF, typename T, int NT, int L, class PROG, int S>
__device__ __forceinline__ void exchange(unsigned char *lds, const regs<T, NT, L, PROG, S> &out,
                                         regs<T, NT, L, PROG, S + 1> &in)
{
    if constexpr (DBUF) {
        static_assert(!SPLIT, "the double-buffered exchange moves whole complex values");
        unsigned char *buf = lds + (S % 2) * PROG::slots() * sizeof(cx<T>);
        write_out<2>(buf, out);
        __syncthreads();
        read_in<2>(buf, in);
        return;
    }
    if constexpr (S > 0) {
        __syncthreads();
    }
    if constexpr (SPLIT) {
        write_out<0>(lds, out);
        __syncthreads();
        read_in<0>(lds, in);
        __syncthreads();
        write_out<1>(lds, out);
        __syncthreads();
        read_in<1>(lds, in);
    } else {
        write_out<2>(lds, out);
        __syncthreads();
        read_in<2>(lds, in);
    }
}

template <bool SPLIT, bool PF, typename T, int NT, int L, int MODE, class PROG, int S>
__device__ __forceinline__ void run_from(const sfh_tile_args &a, const tile_ctx<T> &c, unsigned char *lds,
                                         const cx<T> *tw, regs<T, NT, L, PROG, S> &r,
                                         const twregs<T, NT, L, PROG, S> &w)
{
    if constexpr (S + 1 == PROG::count) {
        compute<PF, T, NT, L, MODE, PROG, S>(a, c, tw, w, r);
        if constexpr ((MODE & SFH_M_PAIR) != 0) {
            store_pair<T, NT, L, MODE, PROG, S>(a, c, lds, r);
        } else if constexpr ((MODE & SFH_M_R2C) != 0) {
            store_r2c<T, NT, L, MODE, PROG, S>(a, c, lds, r);
        } else {
            store<T, NT, L, MODE, PROG, S>(a, c, r);
        }
    } else {
        twregs<T, NT, L, PROG, S + 1> wn;
        if constexpr (PF) {
            load_twiddles(tw, wn); // one step ahead
        }
        compute<PF, T, NT, L, MODE, PROG, S>(a, c, tw, w, r);
        regs<T, NT, L, PROG, S + 1> next;
        exchange<SPLIT, (MODE & SFH_M_DBUF) != 0>(lds, r, next);
        run_from<SPLIT, PF, T, NT, L, MODE, PROG, S + 1>(a, c, lds, tw, next, wn);
    }
}

// Touch this workgroup's share of the plan's table block, one 4-byte load per 128-byte
// line: the lines are split among the workgroups that share an XCD (workgroups are dealt
// to the 8 XCDs round robin, so those of one XCD are g, g + 8, ...), so each XCD's L2 holds
// the block after the first tiles' loads.  Returns a value the caller keeps alive to the
// end of the kernel, so that the loads are not removed.
__device__ __forceinline__ uint32_t warm_tables(const sfh_tile_args &a)
{
    uint32_t acc = 0;
    if (a.warm != nullptr) {
        const uint64_t lines = (a.warm_bytes + 127) / 128;
        const uint64_t groups = gridDim.x / 8 > 0 ? gridDim.x / 8 : 1;
        const uint64_t per = (lines + groups - 1) / groups;
        const uint64_t l0 = (blockIdx.x / 8) % groups * per;
        const uint64_t l1 = l0 + per < lines ? l0 + per : lines;
        const uint32_t *w = static_cast<const uint32_t *>(a.warm);
        for (uint64_t l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
            acc += w[l * 32];
        }
    }
    return acc;
}

} // namespace fx

// Entry: tile g = blockIdx.x is one transform (W = 1) or W columns of a strided axis
// (W > 1), addressed as by the generic kernel.
// WAVES: register budget (waves per SIMD); PAD: the pad layout may be used where linear;
// SPLIT: LDS exchanges move real and imaginary parts separately (half the LDS per tile);
// PF: pass twiddles are loaded one step ahead (needs registers for them across a step).
// With W > 1 the launch has W columns per tile, lanes column-fastest (SFH_WFAST), and each
// column's NT / W work-items run the program on their column, exchanging through the
// column's own LDS region.
template <typename T, int NT, int WAVES, int MODE, bool PAD, bool SPLIT, bool PF, class R1, class R2, int W = 1>
__global__ __launch_bounds__(NT) SFH_WAVES_PER_EU(WAVES)
void fixed_fft(const sfh_tile_args a_byval)
{
    using prog = fx::program<R1, R2, MODE, PAD, W>;
    constexpr int L = prog::L;
    constexpr int NTC = NT / W; // work-items per column
    static_assert(NTC * W == NT, "work-items split evenly over the columns");
    static_assert(W == 1 || (MODE & (SFH_M_PAIR | SFH_M_DBUF | SFH_M_MID)) == 0,
                  "columns per tile: plain and pre/post modes with one LDS buffer");
    static_assert(!(MODE & SFH_M_T4), "no four-step twiddle in the length-specialized kernels");
    static_assert(!(MODE & SFH_M_C2R) || (W == 1 && (MODE & ~(SFH_M_C2R | SFH_M_DBUF)) == 0),
                  "C2R tangle: one contiguous transform per tile, plain mode");
    static_assert(!(MODE & SFH_M_R2C) || (W == 1 && SPLIT && (MODE & ~SFH_M_R2C) == 0),
                  "R2C untangle: one contiguous transform per tile, split exchanges, plain mode");
    static_assert(!(MODE & SFH_M_MID) || (fx::length_of<R2>() == L && R2::v[0] == R1::v[R1::n - 1]),
                  "second FFT: same length, first radix = the first FFT's last");
#if defined(__HIP_DEVICE_COMPILE__)
    const sfh_tile_args &a = *(const sfh_tile_args *)(__builtin_amdgcn_kernarg_segment_ptr());
    (void)a_byval;
#else
    const sfh_tile_args &a = a_byval;
#endif
    SFH_DYNAMIC_LDS(sfh_smem);
    tile_ctx<T> c;
    uint64_t g = blockIdx.x;
    if constexpr (W > 1) {
        // groups of G = 2^lg tiles: g = 8 r + x runs tile G (8 (r / G) + x) + (r mod G), a
        // bijection of the first 8G-multiple of tiles that gives g, g + 8, ..., g + 8 (G - 1)
        // the adjacent tiles G m ... G m + G - 1
        const uint32_t lg = (a.flags & SFH_TILE_XCD_MASK) >> SFH_TILE_XCD_SHIFT;
        if (lg != 0 && g < (a.ntiles & ~((uint64_t(8) << lg) - 1))) {
            const uint64_t x = g & 7u, r = g >> 3;
            g = ((((r >> lg) << 3) + x) << lg) + (r & ((uint64_t(1) << lg) - 1));
        }
    }
    set_tile<T, MODE>(a, c, g, threadIdx.x);
    unsigned char *lds = sfh_smem;
    if constexpr (W > 1) {
        constexpr int elem = SPLIT ? static_cast<int>(sizeof(T)) : static_cast<int>(sizeof(cx<T>));
        lds += static_cast<uint32_t>(threadIdx.x & (W - 1)) * static_cast<uint32_t>(prog::column_pitch(elem) * elem);
    }
    const uint32_t warmed = fx::warm_tables(a);
    fx::regs<T, NTC, L, prog, 0> r;
    const fx::twregs<T, NTC, L, prog, 0> w0{}; // the first step reads HBM: no pass twiddles
    fx::run_from<SPLIT, PF, T, NTC, L, MODE, prog, 0>(a, c, lds, static_cast<const cx<T> *>(a.twtab), r, w0);
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(warmed)); // the touch loads stay (their values are not needed)
#else
    (void)warmed;
#endif
}

} // namespace sfh
