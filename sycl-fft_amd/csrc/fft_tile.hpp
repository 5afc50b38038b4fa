// CDNA4 (gfx950) tile FFT kernels.
//
// One kernel template covers every hot-path transform of the reference's portable
// provider (src/sycl/sycl.cpp:87-339).  Where the reference runs one global pass per
// radix stage plus a gather and a copy-scale pass (14-36 launches per execute), a launch
// here keeps a 2-D tile (W columns x L points) in LDS and runs the whole length-L FFT
// as Stockham radix-{2,3,4,5,7,8,9,16} passes:
//   * pass 0 reads its butterfly inputs straight from HBM (coalesced across lanes),
//   * passes 1..P-1 exchange through LDS (register staging, two barriers per pass),
//   * the last pass writes its butterfly outputs straight to HBM,
// with pre-ops (conjugation, Bluestein chirp, zero padding), an optional pointwise
// product followed by a second FFT (Bluestein convolution), four-step twiddles,
// post-ops (chirp, normalization scale) and masking fused into the same launch.
//
// Only forward butterflies exist: a backward transform is conj(FFT(conj(x))), so the
// direction is a pair of conjugation flags, and the Bluestein inverse FFT reuses the
// forward passes the same way.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_args.h"

#if defined(__HIP__)
#define SFH_WAVES_PER_EU(...) __attribute__((amdgpu_waves_per_eu(__VA_ARGS__)))
#define SFH_DYNAMIC_LDS(name) extern __shared__ __attribute__((aligned(16))) unsigned char name[]
#define SFH_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#else
#define SFH_SCHED_BARRIER()
#define SFH_WAVES_PER_EU(...)
#define SFH_DYNAMIC_LDS(name) unsigned char *name = emu_dynamic_lds()
#endif

// SFH_TW_GLOBAL = 1 builds a variant that reads the pass and four-step twiddle tables (a
// few KiB, L1/L2-resident) from global memory instead of staging them in LDS behind the
// tile, so a tile whose data alone fits a smaller LDS share gets more resident workgroups.
#ifndef SFH_TW_GLOBAL
#define SFH_TW_GLOBAL 0
#endif

namespace sfh
{
#if SFH_TW_GLOBAL
// that variant's instantiations carry their own names, so the two builds of one
// (T, threads, elements) never share a kernel symbol
inline namespace twg
{
#endif

template <typename T>
struct cx
{
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b)
{
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b)
{
    return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s)
{
    return {a.x * s, a.y * s};
}
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a)
{
    return {a.x, -a.y};
}
// multiply by -i
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a)
{
    return {a.y, -a.x};
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, sfh_fastdiv f)
{
    return (__umulhi(n, f.m) + n) >> f.s;
}

// LDS slot of element e of a tile column.  Layout 0: slot e.  Layout 1 (pad): one pad
// slot after every 32 elements, slot e + e/32.  Stockham writes of an even radix at small
// ns go to power-of-two strides (e.g. stride 16 after a radix-16 pass 0), which in layout
// 0 put all 64 lanes of a wave on a few banks; the pad spreads them over every bank.  The
// planner picks the layout per pass boundary such that both sides stay linear in the
// butterfly input/output index t (writer: power-of-two radix and ns; reader: m = L/R a
// multiple of 32 or a power of two), so every LDS access costs one add of a wave-uniform
// offset to a per-butterfly base.
__device__ __forceinline__ uint32_t lds_slot(uint32_t e, uint32_t pad)
{
    return e + (pad ? (e >> 5) : 0u);
}

// ---------------------------------------------------------------------------------
// Forward DFT butterflies, y[k] = sum_t x[t] exp(-2 pi i t k / R), in place on v[0..R).
// ---------------------------------------------------------------------------------

template <typename T, int R>
struct dft;

template <typename T>
struct dft<T, 2>
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <typename T>
struct dft<T, 4>
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        const cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        const cx<T> a2 = v[1] + v[3], a3 = mul_mi(v[1] - v[3]);
        v[0] = a0 + a2;
        v[2] = a0 - a2;
        v[1] = a1 + a3;
        v[3] = a1 - a3;
    }
};

template <typename T>
struct dft<T, 8>
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        const T c = T(0.70710678118654752440);
        const cx<T> a0 = v[0] + v[4], a1 = v[0] - v[4];
        const cx<T> a2 = v[2] + v[6], a3 = mul_mi(v[2] - v[6]);
        const cx<T> b0 = v[1] + v[5], b1 = v[1] - v[5];
        const cx<T> b2 = v[3] + v[7], b3 = mul_mi(v[3] - v[7]);
        const cx<T> e0 = a0 + a2, e2 = a0 - a2, e1 = a1 + a3, e3 = a1 - a3;
        cx<T> o0 = b0 + b2, o2 = b0 - b2, o1 = b1 + b3, o3 = b1 - b3;
        o1 = {c * (o1.x + o1.y), c * (o1.y - o1.x)};
        o2 = mul_mi(o2);
        o3 = {c * (o3.y - o3.x), -c * (o3.x + o3.y)};
        v[0] = e0 + o0;
        v[4] = e0 - o0;
        v[1] = e1 + o1;
        v[5] = e1 - o1;
        v[2] = e2 + o2;
        v[6] = e2 - o2;
        v[3] = e3 + o3;
        v[7] = e3 - o3;
    }
};

template <typename T>
struct dft<T, 16>
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        // 4 x 4 decomposition: n = r + 4 m, k = k1 + 4 k2.
        const T c1 = T(0.92387953251128675613); // cos(pi/8)
        const T s1 = T(0.38268343236508977173); // sin(pi/8)
        const T h = T(0.70710678118654752440);
        cx<T> a[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cx<T> q[4] = {v[r], v[r + 4], v[r + 8], v[r + 12]};
            dft<T, 4>::run(q);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) {
                a[r][k1] = q[k1];
            }
        }
        // twiddles w16^(r k1), forward sign: w16^p = (cos(pi p / 8), -sin(pi p / 8))
        a[1][1] = cmul(a[1][1], cx<T>{c1, -s1});
        a[1][2] = cmul(a[1][2], cx<T>{h, -h});
        a[1][3] = cmul(a[1][3], cx<T>{s1, -c1});
        a[2][1] = cmul(a[2][1], cx<T>{h, -h});
        a[2][2] = mul_mi(a[2][2]);
        a[2][3] = cmul(a[2][3], cx<T>{-h, -h});
        a[3][1] = cmul(a[3][1], cx<T>{s1, -c1});
        a[3][2] = cmul(a[3][2], cx<T>{-h, -h});
        a[3][3] = cmul(a[3][3], cx<T>{-c1, s1});
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cx<T> q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            dft<T, 4>::run(q);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) {
                v[k1 + 4 * k2] = q[k2];
            }
        }
    }
};

// radix 32 as 2 x 16: the even and odd points (n = 2m + r) each take a 16-point DFT, the
// odd ones are twiddled by w32^k, and radix-2 butterflies give k and k + 16
template <typename T>
struct dft<T, 32>
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        const T c1 = T(0.98078528040323044913); // cos(pi/16)
        const T s1 = T(0.19509032201612826785);
        const T c2 = T(0.92387953251128675613); // cos(pi/8)
        const T s2 = T(0.38268343236508977173);
        const T c3 = T(0.83146961230254523708); // cos(3 pi/16)
        const T s3 = T(0.55557023301960222474);
        const T h = T(0.70710678118654752440);
        cx<T> e[16], o[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            e[m] = v[2 * m];
            o[m] = v[2 * m + 1];
        }
        dft<T, 16>::run(e);
        dft<T, 16>::run(o);
        // w32^k = (cos(pi k / 16), -sin(pi k / 16))
        o[1] = cmul(o[1], cx<T>{c1, -s1});
        o[2] = cmul(o[2], cx<T>{c2, -s2});
        o[3] = cmul(o[3], cx<T>{c3, -s3});
        o[4] = cmul(o[4], cx<T>{h, -h});
        o[5] = cmul(o[5], cx<T>{s3, -c3});
        o[6] = cmul(o[6], cx<T>{s2, -c2});
        o[7] = cmul(o[7], cx<T>{s1, -c1});
        o[8] = mul_mi(o[8]);
        o[9] = cmul(o[9], cx<T>{-s1, -c1});
        o[10] = cmul(o[10], cx<T>{-s2, -c2});
        o[11] = cmul(o[11], cx<T>{-s3, -c3});
        o[12] = cmul(o[12], cx<T>{-h, -h});
        o[13] = cmul(o[13], cx<T>{-c3, -s3});
        o[14] = cmul(o[14], cx<T>{-c2, -s2});
        o[15] = cmul(o[15], cx<T>{-c1, -s1});
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            v[k] = e[k] + o[k];
            v[k + 16] = e[k] - o[k];
        }
    }
};

template <typename T>
struct dft<T, 3>;

// radix 9 as 3 x 3: n = r + 3 m, k = k1 + 3 k2, inner twiddles w9^(r k1)
template <typename T>
struct dft9
{
    static __device__ __forceinline__ void run(cx<T> *v);
};

// cos/sin(2 pi p / R) for the odd prime radices
template <int R>
struct odd_consts;
template <>
struct odd_consts<3>
{
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct odd_consts<5>
{
    static constexpr double c[5] = {
        1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
        0.30901699437494742410};
    static constexpr double s[5] = {
        0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
        -0.95105651629515357212};
};
template <>
struct odd_consts<7>
{
    static constexpr double c[7] = {
        1.0,
        0.62348980185873353053,
        -0.22252093395631440429,
        -0.90096886790241912624,
        -0.90096886790241912624,
        -0.22252093395631440429,
        0.62348980185873353053};
    static constexpr double s[7] = {
        0.0,
        0.78183148246802980871,
        0.97492791218182360702,
        0.43388373911755812048,
        -0.43388373911755812048,
        -0.97492791218182360702,
        -0.78183148246802980871};
};

template <typename T, int R>
struct dft_odd
{
    static __device__ __forceinline__ void run(cx<T> *v)
    {
        constexpr int H = (R - 1) / 2;
        cx<T> a[H + 1], b[H + 1];
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            a[m] = v[m] + v[R - m];
            b[m] = v[m] - v[R - m];
        }
        const cx<T> x0 = v[0];
        cx<T> y0 = x0;
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            y0 = y0 + a[m];
        }
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            cx<T> re = x0;
            cx<T> s = {T(0), T(0)};
#pragma unroll
            for (int m = 1; m <= H; ++m) {
                const T cc = T(odd_consts<R>::c[(m * k) % R]);
                const T ss = T(odd_consts<R>::s[(m * k) % R]);
                re.x += cc * a[m].x;
                re.y += cc * a[m].y;
                s.x += ss * b[m].x;
                s.y += ss * b[m].y;
            }
            // y_k = re - i s ; y_{R-k} = re + i s
            v[k] = {re.x + s.y, re.y - s.x};
            v[R - k] = {re.x - s.y, re.y + s.x};
        }
        v[0] = y0;
    }
};

template <typename T>
struct dft<T, 3> : dft_odd<T, 3>
{};
template <typename T>
struct dft<T, 5> : dft_odd<T, 5>
{};
template <typename T>
struct dft<T, 7> : dft_odd<T, 7>
{};

template <typename T>
__device__ __forceinline__ void dft9<T>::run(cx<T> *v)
{
    const T c1 = T(0.76604444311897803520), s1 = T(0.64278760968653932632); // 2 pi / 9
    const T c2 = T(0.17364817766693034885), s2 = T(0.98480775301220805936); // 4 pi / 9
    const T c4 = T(-0.93969262078590838405), s4 = T(0.34202014332566873304); // 8 pi / 9
    cx<T> a[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        cx<T> q[3] = {v[r], v[r + 3], v[r + 6]};
        dft<T, 3>::run(q);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            a[r][k1] = q[k1];
        }
    }
    a[1][1] = cmul(a[1][1], cx<T>{c1, -s1});
    a[1][2] = cmul(a[1][2], cx<T>{c2, -s2});
    a[2][1] = cmul(a[2][1], cx<T>{c2, -s2});
    a[2][2] = cmul(a[2][2], cx<T>{c4, -s4});
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        cx<T> q[3] = {a[0][k1], a[1][k1], a[2][k1]};
        dft<T, 3>::run(q);
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) {
            v[k1 + 3 * k2] = q[k2];
        }
    }
}

template <typename T>
struct dft<T, 9> : dft9<T>
{};


// ---------------------------------------------------------------------------------
// Per-launch context: tile coordinates and lane mapping, shared by every pass.
// ---------------------------------------------------------------------------------

template <typename T>
struct tile_ctx
{
    const cx<T> *tsrc; // tile base in the source
    cx<T> *tdst;       // tile base in the destination
    const cx<T> *src;  // this lane's column in the source
    cx<T> *dst;        // this lane's column in the destination
    uint32_t w0;       // first global column of the tile
    uint32_t wv;       // valid columns in this tile
    uint32_t w;        // this lane's column
    uint32_t lt;       // this lane's position inside its column
    uint32_t blo;
    uint32_t col;      // logical column of this lane's column
    const cx<T> *tw;   // pass-twiddle tables in LDS
    const cx<T> *t4;   // four-step tables in LDS
    bool wvalid;       // this lane's column exists in the tile
    // split-precision modes: this lane's A operands of the 16-point DFT matrix (packed
    // pairs of BF16/FP16 values; words 8h..8h+3 = hi, 8h+4..8h+7 = lo of output half h)
    uint32_t amat[16];
};

__device__ __forceinline__ uint32_t col_of(const sfh_tile_args &a, uint32_t w0, uint32_t blo, uint32_t w)
{
    const uint32_t q = fdiv(w0 + w, a.col_div);
    const uint32_t r = q - fdiv(q, a.col_mod) * a.col_mod.d;
    return r + blo * a.col_bmul;
}

template <typename T, int MODE>
__device__ __forceinline__ cx<T> load_one(const sfh_tile_args &a, const cx<T> *colp, uint32_t j, uint32_t col)
{
    if constexpr (MODE & SFH_M_PRE) {
        const uint32_t p = j * a.pre_jmul + col * a.pre_cmul;
        if (p >= a.nvalid_in) {
            return {T(0), T(0)};
        }
        cx<T> v = colp[static_cast<int32_t>(j) * a.in_j];
        if (a.flags & SFH_CONJ_IN) {
            v.y = -v.y;
        }
        return cmul(v, static_cast<const cx<T> *>(a.pre)[p]);
    } else {
        cx<T> v = colp[static_cast<int32_t>(j) * a.in_j];
        if (a.flags & SFH_CONJ_IN) {
            v.y = -v.y;
        }
        return v;
    }
}

// W_L^m from the two-level pass table in LDS
template <typename T>
__device__ __forceinline__ cx<T> pass_twiddle(const cx<T> *tab, uint32_t bits, uint32_t m)
{
    return cmul(tab[(1u << bits) + (m >> bits)], tab[m & ((1u << bits) - 1u)]);
}

// W_N^m from the three-level four-step table in LDS
template <typename T>
__device__ __forceinline__ cx<T> t4_twiddle(const sfh_tile_args &a, const cx<T> *tab, uint32_t m)
{
    const uint32_t b = a.t4_bits, mask = (1u << b) - 1u;
    const cx<T> w0 = tab[m & mask];
    const cx<T> w1 = tab[(1u << b) + ((m >> b) & mask)];
    const cx<T> w2 = tab[(2u << b) + (m >> (2 * b))];
    return cmul(cmul(w2, w1), w0);
}

// w[t] = W^(e t) for t < R from four table lookups (t = 1, 2, 4, 8) and at most three
// products each: lk(x) returns W^x for x < the table length.
template <typename T, int R, class Lookup>
__device__ __forceinline__ void twiddle_powers(cx<T> *w, uint32_t e, Lookup lk)
{
    w[0] = {T(1), T(0)};
    if constexpr (R > 1) {
        w[1] = lk(e);
    }
    if constexpr (R > 2) {
        w[2] = lk(2 * e);
    }
    if constexpr (R > 4) {
        w[4] = lk(4 * e);
    }
    if constexpr (R > 8) {
        w[8] = lk(8 * e);
    }
#pragma unroll
    for (int t = 3; t < R; ++t) {
        if (t == 4 || t == 8) {
            continue;
        }
        const int hb = t >= 8 ? 8 : (t >= 4 ? 4 : 2);
        w[t] = cmul(w[hb], w[t - hb]);
    }
}

// Store epilogue factors (re * sx, im * sy): conjugations and the normalization scale.
template <typename T>
struct epilogue
{
    T sx, sy;
    __device__ __forceinline__ explicit epilogue(const sfh_tile_args &a)
        : sx(static_cast<T>(a.out_sx)), sy(static_cast<T>(a.out_sy))
    {}
    __device__ __forceinline__ cx<T> operator()(cx<T> v) const
    {
        return {v.x * sx, v.y * sy};
    }
};

// Store pipeline for one element (the coalescing STORE_LDS loop): [T4] -> [POST] -> epilogue.
template <typename T, int MODE>
__device__ __forceinline__ void store_one(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *colp, uint32_t k, uint32_t col, cx<T> v)
{
    if constexpr (MODE & SFH_M_T4) {
        v = cmul(v, t4_twiddle(a, c.t4, col * k));
    }
    if constexpr (MODE & SFH_M_POST) {
        const uint32_t q = k * a.post_kmul + col * a.post_cmul;
        if (q >= a.nvalid_out) {
            return;
        }
        v = cmul(v, static_cast<const cx<T> *>(a.post)[q]);
    }
    colp[static_cast<int32_t>(k) * a.out_k] = epilogue<T>(a)(v);
}

// ---------------------------------------------------------------------------------
// One Stockham pass.  Source/destination of the butterfly data:
//   IN_GLOBAL  : pass 0 reads x[i + t L/R] of its column straight from HBM (pre-ops)
//   IN_LDS     : read element i + t L/R of the LDS column, twiddle by W_{ns R}^{(i mod ns) t}
//   OUT_LDS    : Stockham write of element (i/ns) ns R + i%ns + t ns
//   (element e of column w lives at lds[w LP + lds_slot(e, layout)]; the load loop of
//   LOAD_LDS writes layout 0, which pass 0 then reads)
//   OUT_GLOBAL : last pass (ns R == L, i < ns): output k = i + t L/R goes to HBM
//   OUT_MID    : last pass of FFT #1: spectrum product + conjugation in registers, then
//                pass 0 of FFT #2 (radix R, ns = 1, same input positions), then OUT_LDS
// ---------------------------------------------------------------------------------

enum
{
    IN_GLOBAL = 0,
    IN_LDS = 1,
    IN_REGS = 2 // pass 0 inputs already in the prefetch registers (loaded by prefetch_pass0)
};

// Largest register footprint (butterflies per lane x radix) of any pass radix.
template <int E>
struct prefetch_size
{
    static constexpr int nbr(int r)
    {
        return (E + r - 1) / r * r;
    }
    static constexpr int max2(int x, int y)
    {
        return x > y ? x : y;
    }
    static constexpr int value = max2(
        max2(max2(nbr(2), nbr(3)), max2(nbr(4), nbr(5))),
        max2(max2(nbr(7), nbr(8)), max2(nbr(9), E <= 16 ? nbr(16) : 0)));
};
enum
{
    OUT_LDS = 0,
    OUT_GLOBAL = 1,
    OUT_MID = 2
};

// Load the inputs of pass 0 (radix R) of the lane's butterflies into pf (with pre-ops).
template <typename T, int E, int R, int MODE>
__device__ __forceinline__ void prefetch_pass0(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *pf, bool wvalid)
{
    constexpr int NB = (E + R - 1) / R;
    const uint32_t m = a.L / R;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
        if (wvalid && i < m) {
#pragma unroll
            for (int t = 0; t < R; ++t) {
                pf[q * R + t] = load_one<T, MODE>(a, c.src, i + t * m, c.col);
            }
        }
    }
}

struct no_hook
{
    __device__ void operator()() const {}
};

// ---------------------------------------------------------------------------------
// Split-precision radix-16 passes on MFMA (fp32 plans, modes SFH_M_SPLIT_BF16/_F16).
//
// A radix-16 butterfly is the real 32x32 product [Re y; Im y] = F [Re x; Im x] with the
// 16-point DFT matrix F, re/im interleaved along both dimensions.  A wave computes 16
// butterflies at once with two v_mfma_f32_16x16x32 products (output rows 0..15 = outputs
// 0..7, rows 16..31 = outputs 8..15):
//   A (16 x 32): lane l holds row l%16, columns 8(l/16) .. 8(l/16)+7
//   B (32 x 16): lane l holds column l%16 (its butterfly), rows 8(l/16) .. 8(l/16)+7,
//                i.e. inputs t = 4(l/16) .. 4(l/16)+3 of that butterfly, re/im pairs
//   D (16 x 16): lane l holds column l%16, rows 4(l/16) .. 4(l/16)+3, i.e. outputs
//                2(l/16) and 2(l/16)+1 (and 8 more in the second product)
// so each lane loads, twiddles and stores four of its butterfly's sixteen points, exactly
// the LDS/HBM slots the VALU pass uses for them.  Operands are split x = hi + lo in BF16
// or FP16 and the product is A_lo B_hi + A_hi B_lo + A_hi B_hi (the lo*lo term is dropped),
// accumulated in FP32.  FP16 operands are scaled per butterfly by 2^(15 - e), e the
// binary exponent of the butterfly's largest component, so hi never overflows and lo keeps
// 11 more bits; the outputs are scaled back exactly.
// ---------------------------------------------------------------------------------

constexpr int split_prec(int mode)
{
    return (mode & SFH_M_SPLIT_BF16) ? 1 : ((mode & SFH_M_SPLIT_F16) ? 2 : 0);
}

#if defined(__HIP__)
typedef float sfh_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t sfh_u4 __attribute__((ext_vector_type(4)));
typedef __bf16 sfh_bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 sfh_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 sfh_bf2 __attribute__((ext_vector_type(2)));
typedef _Float16 sfh_h2 __attribute__((ext_vector_type(2)));

// two fp32 values rounded to the operand type (nearest even), element 0 in the low half
template <int P>
__device__ __forceinline__ uint32_t split_pack(float a, float b)
{
    if constexpr (P == 1) {
        const sfh_bf2 v = {static_cast<__bf16>(a), static_cast<__bf16>(b)};
        return __builtin_bit_cast(uint32_t, v);
    } else {
        const sfh_h2 v = {static_cast<_Float16>(a), static_cast<_Float16>(b)};
        return __builtin_bit_cast(uint32_t, v);
    }
}
template <int P>
__device__ __forceinline__ float split_elem0(uint32_t w)
{
    if constexpr (P == 1) {
        return __uint_as_float(w << 16);
    } else {
        return static_cast<float>(__builtin_bit_cast(sfh_h2, w).x);
    }
}
template <int P>
__device__ __forceinline__ float split_elem1(uint32_t w)
{
    if constexpr (P == 1) {
        return __uint_as_float(w & 0xffff0000u);
    } else {
        return static_cast<float>(__builtin_bit_cast(sfh_h2, w).y);
    }
}
// acc += A B for one 16x16x32 block (a, b: this lane's four packed operand words)
template <int P>
__device__ __forceinline__ void mfma16(sfh_f4 &acc, const uint32_t *a, const uint32_t *b)
{
    const sfh_u4 av = {a[0], a[1], a[2], a[3]};
    const sfh_u4 bv = {b[0], b[1], b[2], b[3]};
    if constexpr (P == 1) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(sfh_bf8, av), __builtin_bit_cast(sfh_bf8, bv), acc, 0, 0, 0);
    } else {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(
            __builtin_bit_cast(sfh_h8, av), __builtin_bit_cast(sfh_h8, bv), acc, 0, 0, 0);
    }
}
__device__ __forceinline__ float wave_xor_f(float v, int mask)
{
    return __shfl_xor(v, mask);
}
__device__ __forceinline__ float wave_xor(float v, int mask)
{
    return __shfl_xor(v, mask);
}
__device__ __forceinline__ double wave_xor(double v, int mask)
{
    return __shfl_xor(v, mask);
}
__device__ __forceinline__ int exponent_of(float v)
{
    return __builtin_amdgcn_frexp_expf(v);
}
__device__ __forceinline__ float scale2(float v, int e)
{
    return ldexpf(v, e);
}
__device__ __forceinline__ uint32_t wave_lane()
{
    return __lane_id();
}
#else
// host build (CPU emulation): operands as bit patterns, MFMA and lane exchange emulated
struct sfh_f4
{
    float v[4];
    float &operator[](int i)
    {
        return v[i];
    }
    float operator[](int i) const
    {
        return v[i];
    }
};
template <int P>
inline uint32_t split_pack(float a, float b)
{
    return P == 1 ? (emu_to_bf16(a) | (emu_to_bf16(b) << 16)) : (emu_to_f16(a) | (emu_to_f16(b) << 16));
}
template <int P>
inline float split_elem0(uint32_t w)
{
    return P == 1 ? emu_from_bf16(w & 0xffffu) : emu_from_f16(w & 0xffffu);
}
template <int P>
inline float split_elem1(uint32_t w)
{
    return P == 1 ? emu_from_bf16(w >> 16) : emu_from_f16(w >> 16);
}
template <int P>
inline void mfma16(sfh_f4 &acc, const uint32_t *a, const uint32_t *b)
{
    emu_mfma_16x16x32(P, acc.v, a, b);
}
inline float wave_xor_f(float v, int mask)
{
    return emu_wave_xor(v, mask);
}
inline float wave_xor(float v, int mask)
{
    return emu_wave_xor(v, mask);
}
inline double wave_xor(double v, int mask)
{
    return emu_wave_xor_d(v, mask);
}
inline int exponent_of(float v)
{
    int e = 0;
    (void)std::frexp(v, &e);
    return e;
}
inline float scale2(float v, int e)
{
    return std::ldexp(v, e);
}
inline uint32_t wave_lane()
{
    return threadIdx.x & 63u;
}
#endif

// cos(2 pi p / 16) for any p (bits 0..3 used)
__device__ __forceinline__ float cos16(uint32_t p)
{
    const float k1 = 0.92387953251128675613f; // cos(pi/8)
    const float k2 = 0.70710678118654752440f; // cos(pi/4)
    const float k3 = 0.38268343236508977173f; // cos(3pi/8)
    const uint32_t q = p & 7u;
    const float v = q == 0u ? 1.0f
                  : q == 1u ? k1
                  : q == 2u ? k2
                  : q == 3u ? k3
                  : q == 4u ? 0.0f
                  : q == 5u ? -k3
                  : q == 6u ? -k2
                            : -k1;
    return (p & 8u) ? -v : v;
}

// This lane's A operands of F (forward, y_k = sum_t x_t W16^(k t)): row r = lane%16 of
// output half h is output k = 8h + r/2, real part for even r, imaginary for odd r;
// columns 8g + jj are input t = 4g + jj/2, real part for even jj, imaginary for odd jj.
//   Re y += C Re x + S Im x,   Im y += -S Re x + C Im x,   (C, S) = (cos, sin)(2 pi k t / 16)
template <int P>
__device__ __forceinline__ void split_dft16_matrix(uint32_t *amat, uint32_t lane)
{
    const uint32_t r = lane & 15u, g = lane >> 4;
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t k = 8u * h + (r >> 1);
        float vals[8];
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj) {
            const uint32_t t = 4u * g + (jj >> 1);
            const uint32_t e = (k * t) & 15u;
            const float C = cos16(e), S = cos16(e + 12u); // sin x = cos(x - pi/2)
            vals[jj] = (r & 1u) ? ((jj & 1u) ? C : -S) : ((jj & 1u) ? S : C);
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t hi = split_pack<P>(vals[2 * j], vals[2 * j + 1]);
            amat[8 * h + j] = hi;
            amat[8 * h + 4 + j] =
                split_pack<P>(vals[2 * j] - split_elem0<P>(hi), vals[2 * j + 1] - split_elem1<P>(hi));
        }
    }
}

// Butterfly of a split-precision group: the VALU lane mapping applied to the virtual lane
// vt, so group q of a wave takes the butterflies VALU lanes 16q .. 16q+15 would own.
struct split_bfly
{
    uint32_t i; // butterfly index in its column
    uint32_t w; // column in the tile
    bool ok;    // column exists and i < m
};
__device__ __forceinline__ split_bfly split_butterfly(const sfh_tile_args &a, uint32_t wv, uint32_t vt, uint32_t m)
{
    split_bfly b;
    if (a.W == 1) {
        b.w = 0;
        b.i = vt;
    } else if (a.flags & SFH_WFAST) {
        b.w = vt & (a.W - 1);
        b.i = vt >> a.logW;
    } else {
        b.w = vt >> a.logtpc;
        b.i = vt & (a.tpc - 1);
    }
    b.ok = b.w < wv && b.i < m;
    return b;
}

// One radix-16 Stockham pass of a split-precision launch.  The wave's 64 VALU-mode
// butterflies (one per lane: E = 16 = R) run as four groups of 16: group q takes those of
// VALU lanes 16q .. 16q+15, and lane l works on four of the sixteen points of butterfly l % 16.
template <typename T, int MODE, class Hook>
__device__ __forceinline__ void run_pass16_split(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, const sfh_pass &p, int in_mode,
    int out_mode, Hook after_loads)
{
    static_assert(sizeof(T) == 4, "split-precision passes are for fp32 plans");
    constexpr int P = split_prec(MODE);
    const uint32_t m = a.L >> 4; // butterflies per column
    const uint32_t lane = wave_lane();
    const uint32_t g = lane >> 4;
    const uint32_t vt0 = threadIdx.x - lane + (lane & 15u);
    const bool col_units = a.unit_dim == SFH_UNIT_COL;
    cx<T> v[4][4];

    if (in_mode == IN_GLOBAL) {
        const int64_t tstep = static_cast<int64_t>(m) * a.in_j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const split_bfly b = split_butterfly(a, c.wv, vt0 + 16u * q, m);
            if (b.ok) {
                const cx<T> *src = (col_units && a.in_units)
                                       ? static_cast<const cx<T> *>(a.in_units[c.w0 + b.w])
                                       : c.tsrc + static_cast<int32_t>(b.w) * a.in_w;
                const cx<T> *bq = src + static_cast<int64_t>(b.i) * a.in_j + static_cast<int64_t>(4u * g) * tstep;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[q][j] = bq[j * tstep];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[q][j] = {T(0), T(0)};
                }
            }
        }
        after_loads();
        if (a.flags & SFH_CONJ_IN) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[q][j].y = -v[q][j].y;
                }
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const split_bfly b = split_butterfly(a, c.wv, vt0 + 16u * q, m);
            if (b.ok) {
                const cx<T> *ldsc = lds + b.w * a.LP;
                const uint32_t i = b.i;
                const uint32_t gi = lds_slot(i, p.pad_in);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t tm = (4u * g + static_cast<uint32_t>(j)) * m;
                    v[q][j] = ldsc[gi + tm + (p.pad_in ? (tm >> 5) : 0u)];
                }
                if (p.ns > 1) {
                    // W^(step t) for t = 4g + j: W^(4g step) W^(j step)
                    const uint32_t kq = fdiv(i, p.ns_div);
                    const uint32_t step = (i - kq * p.ns) * p.twstride;
                    const cx<T> w1 = pass_twiddle(c.tw, a.tw_bits, step);
                    const cx<T> w2 = pass_twiddle(c.tw, a.tw_bits, 2u * step);
                    const cx<T> wb = pass_twiddle(c.tw, a.tw_bits, 4u * g * step);
                    const cx<T> wb2 = cmul(wb, w2);
                    v[q][0] = cmul(v[q][0], wb);
                    v[q][1] = cmul(v[q][1], cmul(wb, w1));
                    v[q][2] = cmul(v[q][2], wb2);
                    v[q][3] = cmul(v[q][3], cmul(wb2, w1));
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[q][j] = {T(0), T(0)};
                }
            }
        }
    }

    // 16-point DFTs: two output halves x three split products per group
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float x[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x[2 * j] = v[q][j].x;
            x[2 * j + 1] = v[q][j].y;
        }
        int ex = 0;
        if constexpr (P == 2) {
            float mx = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                mx = fmaxf(mx, fabsf(x[k]));
            }
            mx = fmaxf(mx, wave_xor_f(mx, 16));
            mx = fmaxf(mx, wave_xor_f(mx, 32));
            ex = exponent_of(mx); // mx < 2^ex
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                x[k] = scale2(x[k], 15 - ex);
            }
        }
        uint32_t bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bh[j] = split_pack<P>(x[2 * j], x[2 * j + 1]);
            bl[j] = split_pack<P>(x[2 * j] - split_elem0<P>(bh[j]), x[2 * j + 1] - split_elem1<P>(bh[j]));
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            sfh_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            mfma16<P>(acc, c.amat + 8 * h + 4, bh); // A_lo B_hi
            mfma16<P>(acc, c.amat + 8 * h, bl);     // A_hi B_lo
            mfma16<P>(acc, c.amat + 8 * h, bh);     // A_hi B_hi
            if constexpr (P == 2) {
                v[q][2 * h] = {scale2(acc[0], ex - 15), scale2(acc[1], ex - 15)};
                v[q][2 * h + 1] = {scale2(acc[2], ex - 15), scale2(acc[3], ex - 15)};
            } else {
                v[q][2 * h] = {acc[0], acc[1]};
                v[q][2 * h + 1] = {acc[2], acc[3]};
            }
        }
    }
    // v[q][j] now holds output 2g + (j & 1) + 8 (j >> 1) of group q's butterfly

    if (out_mode == OUT_GLOBAL) {
        const epilogue<T> ep(a);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const split_bfly b = split_butterfly(a, c.wv, vt0 + 16u * q, m);
            if (b.ok) {
                cx<T> *dst = (col_units && a.out_units) ? static_cast<cx<T> *>(a.out_units[c.w0 + b.w])
                                                        : c.tdst + static_cast<int32_t>(b.w) * a.out_w;
                uint32_t col = 0;
                if constexpr (MODE & SFH_M_T4) {
                    col = col_of(a, c.w0, c.blo, b.w);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t k = b.i + (2u * g + (j & 1) + 8u * (j >> 1)) * m;
                    cx<T> y = v[q][j];
                    if constexpr (MODE & SFH_M_T4) {
                        y = cmul(y, t4_twiddle(a, c.t4, col * k));
                    }
                    dst[static_cast<int64_t>(k) * a.out_k] = ep(y);
                }
            }
        }
        return;
    }

    __syncthreads(); // every lane has read this pass's inputs
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const split_bfly b = split_butterfly(a, c.wv, vt0 + 16u * q, m);
        if (b.ok) {
            cx<T> *ldsc = lds + b.w * a.LP;
            const uint32_t kq = fdiv(b.i, p.ns_div);
            const uint32_t gb = lds_slot(kq * p.ns * 15u + b.i, p.pad_out);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t ts = (2u * g + (j & 1) + 8u * (j >> 1)) * p.ns;
                ldsc[gb + ts + (p.pad_out ? (ts >> 5) : 0u)] = v[q][j];
            }
        }
    }
    __syncthreads();
}

// after_loads() runs once every lane has issued its pass-0 HBM loads (IN_GLOBAL), before
// they are consumed: the twiddle tables are staged there so both round trips overlap.
template <typename T, int NT, int E, int R, int MODE, class Hook = no_hook>
__device__ __forceinline__ void run_pass(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, const sfh_pass &p, int in_mode,
    int out_mode, bool wvalid, const cx<T> *pf, Hook after_loads = Hook())
{
    if constexpr (split_prec(MODE) != 0 && R == 16) {
        run_pass16_split<T, MODE>(a, c, lds, p, in_mode, out_mode, after_loads);
        return;
    }
    constexpr int NB = (E + R - 1) / R;
    const uint32_t m = a.L / R; // butterflies per column
    cx<T> *ldsc = lds + c.w * a.LP;
    cx<T> v[NB][R];

    if (in_mode == IN_GLOBAL) {
        // Input t of butterfly i is element i + t m of the column: one 64-bit base per
        // butterfly and a wave-uniform step per t, so an element costs one address add.
        const int64_t tstep = static_cast<int64_t>(m) * a.in_j;
        if constexpr (MODE & SFH_M_PRE) {
            // zero padding: element j exists iff p = j pre_jmul + col pre_cmul < nvalid_in;
            // the chirp value pre[p] is loaded beside it so both round trips overlap
            const cx<T> *pre = static_cast<const cx<T> *>(a.pre);
            const uint32_t pstep = m * a.pre_jmul;
            cx<T> w[NB][R];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
                    const cx<T> *bq = c.src + static_cast<int64_t>(i) * a.in_j;
                    const uint32_t pq = i * a.pre_jmul + c.col * a.pre_cmul;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const uint32_t pp = pq + static_cast<uint32_t>(t) * pstep;
                        if (pp < a.nvalid_in) {
                            v[q][t] = bq[t * tstep];
                            w[q][t] = pre[pp];
                        } else {
                            v[q][t] = {T(0), T(0)};
                            w[q][t] = {T(0), T(0)};
                        }
                    }
                }
            }
            after_loads();
            if (a.flags & SFH_CONJ_IN) { // wave-uniform: a scalar branch around the loop
#pragma unroll
                for (int q = 0; q < NB; ++q) {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        v[q][t].y = -v[q][t].y;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[q][t] = cmul(v[q][t], w[q][t]);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
                    const cx<T> *bq = c.src + static_cast<int64_t>(i) * a.in_j;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        v[q][t] = bq[t * tstep];
                    }
                }
            }
            after_loads();
            if (a.flags & SFH_CONJ_IN) { // wave-uniform: a scalar branch around the loop
#pragma unroll
                for (int q = 0; q < NB; ++q) {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        v[q][t].y = -v[q][t].y;
                    }
                }
            }
        }
    }

#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
        if (wvalid && i < m) {
            if (in_mode == IN_REGS) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[q][t] = pf[q * R + t];
                }
            } else if (in_mode == IN_GLOBAL) {
                // loaded above
            } else {
                // m is a multiple of 32 or a power of two <= 32 whenever pad_in is set, so
                // slot(i + t m) = slot(i) + t m + (t m) / 32 for every i < m
                const uint32_t gi = lds_slot(i, p.pad_in);
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const uint32_t tm = static_cast<uint32_t>(t) * m;
                    v[q][t] = ldsc[gi + tm + (p.pad_in ? (tm >> 5) : 0u)];
                }
                if (p.ns > 1) {
                    const uint32_t kq = fdiv(i, p.ns_div);
                    const uint32_t step = (i - kq * p.ns) * p.twstride;
                    cx<T> w[R];
                    twiddle_powers<T, R>(w, step, [&](uint32_t x) { return pass_twiddle(c.tw, a.tw_bits, x); });
#pragma unroll
                    for (int t = 1; t < R; ++t) {
                        v[q][t] = cmul(v[q][t], w[t]);
                    }
                }
            }
            dft<T, R>::run(v[q]);
        }
        // keep the scheduler from hoisting the next butterfly's loads and twiddles above this
        // one: bounds the live registers to one butterfly's working set plus the outputs
        SFH_SCHED_BARRIER();
    }

    if (out_mode == OUT_GLOBAL) {
        // Output t of butterfly i is element k = i + t m of the column:
        // [T4] -> [POST] -> epilogue -> write, one 64-bit base per butterfly and a
        // wave-uniform step per t.
        if constexpr (MODE & SFH_M_T4) {
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
                    // W_N^(col (i + t m)) = W_N^(col i) (W_N^(col m))^t
                    const auto lk = [&](uint32_t x) { return t4_twiddle(a, c.t4, x); };
                    const cx<T> base = lk(c.col * i);
                    cx<T> w[R];
                    twiddle_powers<T, R>(w, c.col * m, lk);
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        v[q][t] = cmul(v[q][t], cmul(base, w[t]));
                    }
                }
                SFH_SCHED_BARRIER();
            }
        }
        const epilogue<T> ep(a);
        const int64_t tstep = static_cast<int64_t>(m) * a.out_k;
        // POST: output k exists iff pq = k post_kmul + col post_cmul < nvalid_out and is
        // multiplied by post[pq]
        const uint32_t qstep = m * a.post_kmul;
        const cx<T> *post = static_cast<const cx<T> *>(a.post);
        if constexpr (MODE & SFH_M_POST) {
            // every post-table load before the first store (one in-order counter for vector
            // loads and stores: a load after a store would also wait for the store)
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
                    const uint32_t qq = i * a.post_kmul + c.col * a.post_cmul;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const uint32_t pq = qq + static_cast<uint32_t>(t) * qstep;
                        if (pq < a.nvalid_out) {
                            v[q][t] = cmul(v[q][t], post[pq]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
            if (wvalid && i < m) {
                cx<T> *bq = c.dst + static_cast<int64_t>(i) * a.out_k;
                if constexpr (MODE & SFH_M_POST) {
                    const uint32_t qq = i * a.post_kmul + c.col * a.post_cmul;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const uint32_t pq = qq + static_cast<uint32_t>(t) * qstep;
                        if (pq < a.nvalid_out) {
                            bq[t * tstep] = ep(v[q][t]);
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        bq[t * tstep] = ep(v[q][t]);
                    }
                }
            }
            SFH_SCHED_BARRIER();
        }
        return;
    }

    if constexpr (MODE & SFH_M_MID) {
        if (out_mode == OUT_MID) {
            const cx<T> *mid = static_cast<const cx<T> *>(a.mid);
            const uint32_t sstep = m * a.mid_kmul;
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
                    const uint32_t sq = i * a.mid_kmul + c.col * a.mid_cmul;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        v[q][t] = cconj(cmul(v[q][t], mid[sq + static_cast<uint32_t>(t) * sstep]));
                    }
                    dft<T, R>::run(v[q]);
                }
                SFH_SCHED_BARRIER();
            }
        }
    }

    __syncthreads(); // every lane has read this pass's inputs
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
        if (wvalid && i < m) {
            uint32_t base, stride;
            if ((MODE & SFH_M_MID) && out_mode == OUT_MID) {
                base = i * R; // FFT #2 pass 0: ns = 1
                stride = 1;
            } else {
                const uint32_t kq = fdiv(i, p.ns_div);
                base = kq * p.ns * (R - 1) + i; // kq ns R + (i - kq ns)
                stride = p.ns;
            }
            // pad_out is set only for power-of-two R and ns, where element base + t ns stays
            // inside base's 32-block run: slot(base + t ns) = slot(base) + t ns + (t ns) / 32
            const uint32_t gb = lds_slot(base, p.pad_out);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t ts = static_cast<uint32_t>(t) * stride;
                ldsc[gb + ts + (p.pad_out ? (ts >> 5) : 0u)] = v[q][t];
            }
        }
    }
    __syncthreads();
}

template <typename T, int E, int MODE>
__device__ __forceinline__ void dispatch_prefetch(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *pf, bool wvalid)
{
    switch (a.pass[0].radix) {
    case 2:
        prefetch_pass0<T, E, 2, MODE>(a, c, pf, wvalid);
        break;
    case 3:
        prefetch_pass0<T, E, 3, MODE>(a, c, pf, wvalid);
        break;
    case 4:
        prefetch_pass0<T, E, 4, MODE>(a, c, pf, wvalid);
        break;
    case 5:
        prefetch_pass0<T, E, 5, MODE>(a, c, pf, wvalid);
        break;
    case 7:
        prefetch_pass0<T, E, 7, MODE>(a, c, pf, wvalid);
        break;
    case 8:
        prefetch_pass0<T, E, 8, MODE>(a, c, pf, wvalid);
        break;
    case 9:
        prefetch_pass0<T, E, 9, MODE>(a, c, pf, wvalid);
        break;
    default:
        if constexpr (E <= 16) {
            prefetch_pass0<T, E, 16, MODE>(a, c, pf, wvalid);
        }
        break;
    }
}

template <typename T, int NT, int E, int MODE, class Hook = no_hook>
__device__ __forceinline__ void dispatch_pass(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, const sfh_pass &p, int in_mode,
    int out_mode, bool wvalid, const cx<T> *pf, Hook hook = Hook())
{
    switch (p.radix) {
    case 2:
        run_pass<T, NT, E, 2, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 3:
        run_pass<T, NT, E, 3, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 4:
        run_pass<T, NT, E, 4, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 5:
        run_pass<T, NT, E, 5, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 7:
        run_pass<T, NT, E, 7, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 8:
        run_pass<T, NT, E, 8, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    case 9:
        run_pass<T, NT, E, 9, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        break;
    default:
        // radix 16 only where at most 16 elements per lane are staged (register budget);
        // the planner never gives larger-E variants a radix above 9
        if constexpr (E <= 16) {
            run_pass<T, NT, E, 16, MODE>(a, c, lds, p, in_mode, out_mode, wvalid, pf, hook);
        }
        break;
    }
}

// Out mode of fused-program step s (FFT #1 passes 0..np-1, then FFT #2 passes 1..np-1).
__device__ __forceinline__ int out_mode_of(uint32_t s, uint32_t total, uint32_t np, bool two, uint32_t flags)
{
    if (s + 1 == total) {
        return (flags & SFH_STORE_LDS) ? OUT_LDS : OUT_GLOBAL;
    }
    if (two && s + 1 == np) {
        return OUT_MID;
    }
    return OUT_LDS;
}

// Run the `count` passes of radix R (inputs in LDS) that come next in the fused program.
template <typename T, int NT, int E, int R, int MODE>
__device__ __forceinline__ void run_group(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, uint32_t count, uint32_t &s,
    uint32_t total, uint32_t np)
{
    constexpr bool two = (MODE & SFH_M_MID) != 0;
    for (uint32_t k = 0; k < count; ++k, ++s) {
        const uint32_t pi = (s < np) ? s : s + 1;
        run_pass<T, NT, E, R, MODE>(a, c, lds, a.pass[pi], IN_LDS, out_mode_of(s, total, np, two, a.flags),
                                    c.wvalid, nullptr);
    }
}

// The whole fused program of one tile: FFT #1 radix groups in kernel order, then (MID
// modes) FFT #2 groups in reverse order, whose first pass is fused into FFT #1's last.
template <typename T, int NT, int E, int MODE, class After0, class Hook = no_hook>
__device__ __forceinline__ void run_program(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, int first_in, const cx<T> *pf,
    After0 after0, Hook hook = Hook())
{
    constexpr bool two = (MODE & SFH_M_MID) != 0;
    const uint32_t np = a.npass;
    const uint32_t total = two ? 2 * np - 1 : np;
    // pass 0 on its own (its inputs come from registers or HBM), so the pass-0 register
    // buffer is dead before the radix groups run
    dispatch_pass<T, NT, E, MODE>(a, c, lds, a.pass[0], first_in, out_mode_of(0, total, np, two, a.flags),
                                  c.wvalid, pf, hook);
    after0();
    const uint8_t *n = a.rcount;
    uint32_t s = 1;
    if constexpr (E <= 16) {
        run_group<T, NT, E, 16, MODE>(a, c, lds, n[0], s, total, np);
    }
    run_group<T, NT, E, 8, MODE>(a, c, lds, n[1], s, total, np);
    run_group<T, NT, E, 4, MODE>(a, c, lds, n[2], s, total, np);
    run_group<T, NT, E, 2, MODE>(a, c, lds, n[3], s, total, np);
    run_group<T, NT, E, 9, MODE>(a, c, lds, n[4], s, total, np);
    run_group<T, NT, E, 3, MODE>(a, c, lds, n[5], s, total, np);
    run_group<T, NT, E, 5, MODE>(a, c, lds, n[6], s, total, np);
    run_group<T, NT, E, 7, MODE>(a, c, lds, n[7], s, total, np);
    if constexpr (two) {
        // FFT #2: reverse order; its first pass (radix of FFT #1's last) already ran fused
        const uint8_t *m = a.rcount2;
        run_group<T, NT, E, 7, MODE>(a, c, lds, m[7], s, total, np);
        run_group<T, NT, E, 5, MODE>(a, c, lds, m[6], s, total, np);
        run_group<T, NT, E, 3, MODE>(a, c, lds, m[5], s, total, np);
        run_group<T, NT, E, 9, MODE>(a, c, lds, m[4], s, total, np);
        run_group<T, NT, E, 2, MODE>(a, c, lds, m[3], s, total, np);
        run_group<T, NT, E, 4, MODE>(a, c, lds, m[2], s, total, np);
        run_group<T, NT, E, 8, MODE>(a, c, lds, m[1], s, total, np);
        if constexpr (E <= 16) {
            run_group<T, NT, E, 16, MODE>(a, c, lds, m[0], s, total, np);
        }
    }
}

// Coalescing loop HBM -> LDS (natural order), used when the lane mapping of pass 0
// cannot read the source contiguously (e.g. rows in, columns out).
template <typename T, int NT, int E, int MODE>
__device__ __forceinline__ void load_via_lds(const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds)
{
    const bool wcontig = (a.in_w == 1) && a.W > 1;
#pragma unroll
    for (int n = 0; n < E; ++n) {
        const uint32_t e = threadIdx.x + static_cast<uint32_t>(n) * NT;
        uint32_t w, j;
        if (wcontig) {
            w = e & (a.W - 1);
            j = e >> a.logW;
        } else {
            w = fdiv(e, a.L_div);
            j = e - w * a.L;
        }
        if (j < a.L && w < c.wv) {
            const uint32_t col = (MODE & SFH_M_PRE) ? col_of(a, c.w0, c.blo, w) : 0u;
            lds[w * a.LP + j] =
                load_one<T, MODE>(a, c.tsrc + static_cast<int32_t>(w) * a.in_w, j, col);
        }
    }
    __syncthreads();
}

template <typename T, int NT, int E, int MODE>
__device__ __forceinline__ void store_via_lds(const sfh_tile_args &a, const tile_ctx<T> &c, const cx<T> *lds)
{
    const bool wcontig = (a.out_w == 1) && a.W > 1;
    if constexpr (MODE & SFH_M_POST) {
        // every post-table load before the first store: vector loads and stores share one
        // in-order completion counter, so a load after a store would wait for the store
        cx<T> pv[E];
        const cx<T> *post = static_cast<const cx<T> *>(a.post);
#pragma unroll
        for (int n = 0; n < E; ++n) {
            const uint32_t e = threadIdx.x + static_cast<uint32_t>(n) * NT;
            uint32_t w, k;
            if (wcontig) {
                w = e & (a.W - 1);
                k = e >> a.logW;
            } else {
                w = fdiv(e, a.L_div);
                k = e - w * a.L;
            }
            pv[n] = {T(0), T(0)};
            if (k < a.L && w < c.wv) {
                const uint32_t q = k * a.post_kmul + col_of(a, c.w0, c.blo, w) * a.post_cmul;
                if (q < a.nvalid_out) {
                    pv[n] = post[q];
                }
            }
        }
#pragma unroll
        for (int n = 0; n < E; ++n) {
            const uint32_t e = threadIdx.x + static_cast<uint32_t>(n) * NT;
            uint32_t w, k;
            if (wcontig) {
                w = e & (a.W - 1);
                k = e >> a.logW;
            } else {
                w = fdiv(e, a.L_div);
                k = e - w * a.L;
            }
            if (k < a.L && w < c.wv) {
                const uint32_t col = col_of(a, c.w0, c.blo, w);
                if (k * a.post_kmul + col * a.post_cmul < a.nvalid_out) {
                    cx<T> v = lds[w * a.LP + lds_slot(k, a.pad_store)];
                    if constexpr (MODE & SFH_M_T4) {
                        v = cmul(v, t4_twiddle(a, c.t4, col * k));
                    }
                    cx<T> *colp = c.tdst + static_cast<int32_t>(w) * a.out_w;
                    colp[static_cast<int32_t>(k) * a.out_k] = epilogue<T>(a)(cmul(v, pv[n]));
                }
            }
        }
        return;
    }
#pragma unroll
    for (int n = 0; n < E; ++n) {
        const uint32_t e = threadIdx.x + static_cast<uint32_t>(n) * NT;
        uint32_t w, k;
        if (wcontig) {
            w = e & (a.W - 1);
            k = e >> a.logW;
        } else {
            w = fdiv(e, a.L_div);
            k = e - w * a.L;
        }
        if (k < a.L && w < c.wv) {
            const uint32_t col =
                (MODE & (SFH_M_T4 | SFH_M_POST)) ? col_of(a, c.w0, c.blo, w) : 0u;
            store_one<T, MODE>(
                a, c, c.tdst + static_cast<int32_t>(w) * a.out_w, k, col, lds[w * a.LP + lds_slot(k, a.pad_store)]);
        }
    }
}

// Coordinates of tile g and this lane's column in it.
template <typename T, int MODE>
__device__ __forceinline__ void set_tile(const sfh_tile_args &a, tile_ctx<T> &c, uint64_t g, uint32_t tid)
{
    if (a.flags & SFH_TILE_REVERSE) {
        g = a.ntiles - 1 - g;
    }
    const uint64_t wt = g % a.nwt;
    const uint64_t rest = g / a.nwt;
    const uint64_t blo = rest % a.nblo;
    const uint64_t bhi = rest / a.nblo;
    const uint64_t w0 = wt * a.W;
    c.w0 = static_cast<uint32_t>(w0);
    c.blo = static_cast<uint32_t>(blo);
    const uint64_t left = a.wcount - w0;
    c.wv = static_cast<uint32_t>(left < a.W ? left : a.W);
    // Unit tables replace in/out + (index along unit_dim) * stride by a per-unit base.  The
    // BLO/BHI cases are uniform over the workgroup; the COL case is per column and is
    // resolved per lane below (such launches never use the LDS load/store loops).
    const uint32_t ud = a.unit_dim;
    if (a.in_units && ud != SFH_UNIT_COL) {
        const cx<T> *u = static_cast<const cx<T> *>(a.in_units[ud == SFH_UNIT_BLO ? blo : bhi]);
        c.tsrc = u + (ud == SFH_UNIT_BLO ? static_cast<int64_t>(bhi) * a.in_bhi : static_cast<int64_t>(blo) * a.in_blo)
                 + static_cast<int64_t>(w0) * a.in_w;
    } else {
        c.tsrc = static_cast<const cx<T> *>(a.in)
                 + (static_cast<int64_t>(bhi) * a.in_bhi + static_cast<int64_t>(blo) * a.in_blo
                    + static_cast<int64_t>(w0) * a.in_w);
    }
    if (a.out_units && ud != SFH_UNIT_COL) {
        cx<T> *u = static_cast<cx<T> *>(a.out_units[ud == SFH_UNIT_BLO ? blo : bhi]);
        c.tdst = u + (ud == SFH_UNIT_BLO ? static_cast<int64_t>(bhi) * a.out_bhi : static_cast<int64_t>(blo) * a.out_blo)
                 + static_cast<int64_t>(w0) * a.out_w;
    } else {
        c.tdst = static_cast<cx<T> *>(a.out)
                 + (static_cast<int64_t>(bhi) * a.out_bhi + static_cast<int64_t>(blo) * a.out_blo
                    + static_cast<int64_t>(w0) * a.out_w);
    }
    if (a.W == 1) {
        c.w = 0; // NT need not be a power of two here
        c.lt = tid;
    } else if (a.flags & SFH_WFAST) {
        c.w = tid & (a.W - 1);
        c.lt = tid >> a.logW;
    } else {
        c.w = tid >> a.logtpc;
        c.lt = tid & (a.tpc - 1);
    }
    c.wvalid = c.w < c.wv;
    c.src = c.tsrc + static_cast<int32_t>(c.w) * a.in_w;
    c.dst = c.tdst + static_cast<int32_t>(c.w) * a.out_w;
    if (ud == SFH_UNIT_COL) {
        const uint32_t col = c.wvalid ? c.w0 + c.w : c.w0; // invalid lanes never access memory
        if (a.in_units) {
            c.src = static_cast<const cx<T> *>(a.in_units[col]);
        }
        if (a.out_units) {
            c.dst = static_cast<cx<T> *>(a.out_units[col]);
        }
    }
    c.col = (MODE & (SFH_M_PRE | SFH_M_MID | SFH_M_T4 | SFH_M_POST)) ? col_of(a, c.w0, c.blo, c.w) : 0u;
}

// Waves per SIMD that the LDS footprint allows; used as the register-budget hint.
template <typename T, int NT, int E>
struct occupancy
{
    static constexpr int CAP = NT * E;
    static constexpr int BYTES = CAP * int(sizeof(cx<T>)) + 1024;
    static constexpr int WG = 163840 / BYTES;
    static constexpr int RAW = WG * (NT / 64) / 4;
#ifdef SFH_WAVES_EU
    // set per variant where its tiles are sized below CAP (register budget of that many waves)
    static constexpr int WAVES = SFH_WAVES_EU;
#else
    static constexpr int WAVES = RAW < 1 ? 1 : (RAW > 8 ? 8 : RAW);
#endif
};

// Where a tile's twiddle tables live: staged into LDS at a.lds_tab by stage_tables (which
// ends with the barrier that publishes them), or, in the SFH_TW_GLOBAL build, read in place
// from the plan's table block (stage_tables then only keeps the barrier).
template <typename T, int NT>
__device__ __forceinline__ void bind_tables(const sfh_tile_args &a, cx<T> *lds, tile_ctx<T> &c)
{
#if SFH_TW_GLOBAL
    (void)lds;
    const cx<T> *const src = static_cast<const cx<T> *>(a.twtab);
    c.tw = src;
    c.t4 = src + a.tw_entries;
#else
    c.tw = lds + a.lds_tab;
    c.t4 = lds + a.lds_tab + a.tw_entries;
#endif
}

template <typename T, int NT>
__device__ __forceinline__ void stage_tables(const sfh_tile_args &a, cx<T> *lds, uint32_t tid)
{
#if !SFH_TW_GLOBAL
    cx<T> *const tab = lds + a.lds_tab;
    const cx<T> *const tab_src = static_cast<const cx<T> *>(a.twtab);
    const uint32_t tab_count = a.tw_entries + a.t4_entries;
    for (uint32_t e = tid; e < tab_count; e += NT) {
        tab[e] = tab_src[e];
    }
#else
    (void)a;
    (void)lds;
    (void)tid;
#endif
    __syncthreads();
}

// One tile g of a launch: stage the twiddle tables, load, run the fused program, store.
template <typename T, int NT, int E, int MODE>
__device__ __forceinline__ void one_tile(const sfh_tile_args &a, cx<T> *lds, uint64_t g, uint32_t tid)
{
    tile_ctx<T> c;
    if constexpr (split_prec(MODE) != 0) {
        split_dft16_matrix<split_prec(MODE)>(c.amat, wave_lane());
    }
    bind_tables<T, NT>(a, lds, c);
    // Twiddle tables (a few KiB) go to LDS behind the tile (except in the SFH_TW_GLOBAL
    // build).  They are loaded right after the tile's pass-0 loads have been issued, so the
    // two HBM/L2 round trips overlap instead of running one after the other.
    const auto finish_tables = [=]() { stage_tables<T, NT>(a, lds, tid); };
    set_tile<T, MODE>(a, c, g, tid);
    if (a.flags & SFH_LOAD_LDS) {
        finish_tables();
        load_via_lds<T, NT, E, MODE>(a, c, lds);
        run_program<T, NT, E, MODE>(a, c, lds, IN_LDS, nullptr, [] {});
    } else {
        run_program<T, NT, E, MODE>(a, c, lds, IN_GLOBAL, nullptr, [] {}, finish_tables);
    }
    if (a.flags & SFH_STORE_LDS) {
        store_via_lds<T, NT, E, MODE>(a, c, lds);
    }
}

// PREF = 1: persistent workgroups with register prefetch of the next tile's pass-0 inputs
// (launched with at most the resident number of workgroups); PREF = 0: pass 0 reads HBM
// directly (one tile per workgroup, full occupancy).
template <typename T, int NT, int E, int MODE, int PREF>
__global__ __launch_bounds__(NT) SFH_WAVES_PER_EU(PREF ? (occupancy<T, NT, E>::WAVES + 1) / 2 : occupancy<T, NT, E>::WAVES)
void tile_fft(const sfh_tile_args a_byval)
{
    // Read the arguments in place in the kernarg segment (constant address space, scalar
    // loads): indexing a by-value aggregate dynamically would copy it to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
    const sfh_tile_args &a = *(const sfh_tile_args *)(__builtin_amdgcn_kernarg_segment_ptr());
    (void)a_byval;
#else
    const sfh_tile_args &a = a_byval;
#endif
    SFH_DYNAMIC_LDS(sfh_smem);
    cx<T> *lds = reinterpret_cast<cx<T> *>(sfh_smem);
    const uint32_t tid = threadIdx.x;

    if constexpr (!PREF) {
        one_tile<T, NT, E, MODE>(a, lds, blockIdx.x, tid);
        return;
    }
    tile_ctx<T> c;
    if constexpr (split_prec(MODE) != 0) {
        split_dft16_matrix<split_prec(MODE)>(c.amat, wave_lane());
    }
    bind_tables<T, NT>(a, lds, c);
    const uint32_t f = a.flags;
    const auto finish_tables = [=]() { stage_tables<T, NT>(a, lds, tid); };
    if constexpr (PREF) {
        // Persistent loop over tiles g = blockIdx.x, + gridDim.x, ...  The inputs of the next
        // tile's pass 0 are loaded into registers right after the current tile's pass 0, so
        // the HBM reads overlap the LDS passes (no pass waits on global memory).
        uint64_t g = blockIdx.x;
        if (g >= a.ntiles) {
            return; // uniform across the workgroup
        }
        set_tile<T, MODE>(a, c, g, tid);
        cx<T> pf[prefetch_size<E>::value];
        dispatch_prefetch<T, E, MODE>(a, c, pf, c.wvalid);
        finish_tables();
        for (;;) {
            const uint64_t gn = g + gridDim.x;
            const bool more = gn < a.ntiles;
            tile_ctx<T> cn = c;
            run_program<T, NT, E, MODE>(a, c, lds, IN_REGS, pf, [&] {
                if (more) {
                    set_tile<T, MODE>(a, cn, gn, tid);
                    dispatch_prefetch<T, E, MODE>(a, cn, pf, cn.wvalid);
                }
            });
            if (f & SFH_STORE_LDS) {
                store_via_lds<T, NT, E, MODE>(a, c, lds);
            }
            if (!more) {
                break;
            }
            __syncthreads(); // the LDS tile is reused by the next tile
            g = gn;
            c = cn;
        }
    }
}

// Work-queue form: persistent workgroups take items (tile of some launch) from a device
// counter until every item is taken; each item runs with its own launch's arguments, so
// single-launch plans of any length, batch count, direction and normalization that share
// this variant run in one dispatch.  Every workgroup leaves once the counter passes nitems.
template <typename T, int NT, int E, int MODE>
__global__ __launch_bounds__(NT) SFH_WAVES_PER_EU(occupancy<T, NT, E>::WAVES)
void tile_fft_queue(const sfh_queue_args q)
{
    SFH_DYNAMIC_LDS(sfh_smem);
    cx<T> *lds = reinterpret_cast<cx<T> *>(sfh_smem);
    uint32_t *next = reinterpret_cast<uint32_t *>(sfh_smem + q.lds_next);
    const uint32_t tid = threadIdx.x;
    for (;;) {
        // the barrier also ends the previous item's last LDS reads before the next item's writes
        if (tid == 0) {
            *next = atomicAdd(q.counter, 1u);
        }
        __syncthreads();
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t k = __builtin_amdgcn_readfirstlane(*next);
#else
        const uint32_t k = *next;
#endif
        __syncthreads(); // every lane has read the slot before lane 0 takes the next item
        if (k >= q.nitems) {
            return;
        }
        const sfh_queue_item it = q.items[k];
        one_tile<T, NT, E, MODE>(q.args[it.arg], lds, it.tile, tid);
    }
}

#if SFH_TW_GLOBAL
} // namespace twg
#endif
} // namespace sfh
