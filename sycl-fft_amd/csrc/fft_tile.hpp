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
#else
#define SFH_WAVES_PER_EU(...)
#endif

namespace sfh
{

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

template <typename T, int MODE>
__device__ __forceinline__ void store_one(
    const sfh_tile_args &a, cx<T> *colp, uint32_t k, uint32_t col, cx<T> v)
{
    const uint32_t f = a.flags;
    if constexpr (MODE & SFH_M_T4) {
        const uint32_t m = col * k;
        const cx<T> lo = static_cast<const cx<T> *>(a.t4lo)[m & ((1u << a.t4bits) - 1u)];
        const cx<T> hi = static_cast<const cx<T> *>(a.t4hi)[m >> a.t4bits];
        v = cmul(v, cmul(hi, lo));
    }
    if (f & SFH_CONJ_POST) {
        v.y = -v.y;
    }
    if constexpr (MODE & SFH_M_POST) {
        const uint32_t q = k * a.post_kmul + col * a.post_cmul;
        if (q >= a.nvalid_out) {
            return;
        }
        v = cmul(v, static_cast<const cx<T> *>(a.post)[q]);
    }
    if (f & SFH_SCALE) {
        v = cscale(v, static_cast<T>(a.scale));
    }
    if (f & SFH_CONJ_OUT) {
        v.y = -v.y;
    }
    colp[static_cast<int32_t>(k) * a.out_k] = v;
}

// ---------------------------------------------------------------------------------
// One Stockham pass.  Source/destination of the butterfly data:
//   IN_GLOBAL  : pass 0 reads x[i + t L/R] of its column straight from HBM (pre-ops)
//   IN_LDS     : read lds[w LP + i + t L/R], twiddle by W_{ns R}^{(i mod ns) t}
//   OUT_LDS    : Stockham write lds[w LP + (i/ns) ns R + i%ns + t ns]
//   OUT_GLOBAL : last pass (ns R == L, i < ns): output k = i + t L/R goes to HBM
//   OUT_MID    : last pass of FFT #1: spectrum product + conjugation in registers, then
//                pass 0 of FFT #2 (radix R, ns = 1, same input positions), then OUT_LDS
// ---------------------------------------------------------------------------------

enum
{
    IN_GLOBAL = 0,
    IN_LDS = 1
};
enum
{
    OUT_LDS = 0,
    OUT_GLOBAL = 1,
    OUT_MID = 2
};

template <typename T, int NT, int E, int R, int MODE>
__device__ __forceinline__ void run_pass(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, const sfh_pass &p, int in_mode,
    int out_mode, bool wvalid)
{
    constexpr int NB = (E + R - 1) / R;
    const uint32_t m = a.L / R; // butterflies per column
    const cx<T> *tw = static_cast<const cx<T> *>(a.tw);
    cx<T> *ldsc = lds + c.w * a.LP;
    cx<T> v[NB][R];

#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
        if (wvalid && i < m) {
            if (in_mode == IN_GLOBAL) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[q][t] = load_one<T, MODE>(a, c.src, i + t * m, c.col);
                }
            } else {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    v[q][t] = ldsc[i + t * m];
                }
                if (p.ns > 1) {
                    const uint32_t kq = fdiv(i, p.ns_div);
                    const uint32_t step = (i - kq * p.ns) * p.twstride;
#pragma unroll
                    for (int t = 1; t < R; ++t) {
                        v[q][t] = cmul(v[q][t], tw[step * t]);
                    }
                }
            }
            dft<T, R>::run(v[q]);
        }
    }

    if (out_mode == OUT_GLOBAL) {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
            if (wvalid && i < m) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    store_one<T, MODE>(a, c.dst, i + t * m, c.col, v[q][t]);
                }
            }
        }
        return;
    }

    if constexpr (MODE & SFH_M_MID) {
        if (out_mode == OUT_MID) {
            const cx<T> *mid = static_cast<const cx<T> *>(a.mid);
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint32_t i = c.lt + static_cast<uint32_t>(q) * a.tpc;
                if (wvalid && i < m) {
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const uint32_t s = (i + t * m) * a.mid_kmul + c.col * a.mid_cmul;
                        v[q][t] = cconj(cmul(v[q][t], mid[s]));
                    }
                    dft<T, R>::run(v[q]);
                }
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
#pragma unroll
            for (int t = 0; t < R; ++t) {
                ldsc[base + t * stride] = v[q][t];
            }
        }
    }
    __syncthreads();
}

template <typename T, int NT, int E, int MODE>
__device__ __forceinline__ void dispatch_pass(
    const sfh_tile_args &a, const tile_ctx<T> &c, cx<T> *lds, const sfh_pass &p, int in_mode,
    int out_mode, bool wvalid)
{
    switch (p.radix) {
    case 2:
        run_pass<T, NT, E, 2, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 3:
        run_pass<T, NT, E, 3, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 4:
        run_pass<T, NT, E, 4, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 5:
        run_pass<T, NT, E, 5, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 7:
        run_pass<T, NT, E, 7, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 8:
        run_pass<T, NT, E, 8, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    case 9:
        run_pass<T, NT, E, 9, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        break;
    default:
        // radix 16 only where at most 16 elements per lane are staged (register budget);
        // the planner never gives larger-E variants a radix above 9
        if constexpr (E <= 16) {
            run_pass<T, NT, E, 16, MODE>(a, c, lds, p, in_mode, out_mode, wvalid);
        }
        break;
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
                a, c.tdst + static_cast<int32_t>(w) * a.out_w, k, col, lds[w * a.LP + k]);
        }
    }
}

// Waves per SIMD that the LDS footprint allows; used as the register-budget hint.
template <typename T, int NT, int E>
struct occupancy
{
    static constexpr int CAP = NT * E;
    static constexpr int SLACK = ((CAP + 64) * int(sizeof(cx<T>)) > 163840) ? 0 : 64;
    static constexpr int BYTES = (CAP + SLACK) * int(sizeof(cx<T>));
    static constexpr int WG = 163840 / BYTES;
    static constexpr int RAW = WG * (NT / 64) / 4;
    static constexpr int WAVES = RAW < 1 ? 1 : (RAW > 8 ? 8 : RAW);
};

template <typename T, int NT, int E, int MODE>
__global__ __launch_bounds__(NT) SFH_WAVES_PER_EU(occupancy<T, NT, E>::WAVES)
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
    __shared__ cx<T> lds[occupancy<T, NT, E>::CAP + occupancy<T, NT, E>::SLACK];

    tile_ctx<T> c;
    {
        const uint64_t g = blockIdx.x;
        const uint64_t wt = g % a.nwt;
        const uint64_t rest = g / a.nwt;
        const uint64_t blo = rest % a.nblo;
        const uint64_t bhi = rest / a.nblo;
        const uint64_t w0 = wt * a.W;
        c.w0 = static_cast<uint32_t>(w0);
        c.blo = static_cast<uint32_t>(blo);
        const uint64_t left = a.wcount - w0;
        c.wv = static_cast<uint32_t>(left < a.W ? left : a.W);
        c.tsrc = static_cast<const cx<T> *>(a.in)
                 + (static_cast<int64_t>(bhi) * a.in_bhi + static_cast<int64_t>(blo) * a.in_blo
                    + static_cast<int64_t>(w0) * a.in_w);
        c.tdst = static_cast<cx<T> *>(a.out)
                 + (static_cast<int64_t>(bhi) * a.out_bhi + static_cast<int64_t>(blo) * a.out_blo
                    + static_cast<int64_t>(w0) * a.out_w);
    }
    const uint32_t tid = threadIdx.x;
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
    const bool wvalid = c.w < c.wv;
    c.src = c.tsrc + static_cast<int32_t>(c.w) * a.in_w;
    c.dst = c.tdst + static_cast<int32_t>(c.w) * a.out_w;
    c.col = (MODE != 0) ? col_of(a, c.w0, c.blo, c.w) : 0u;

    const uint32_t f = a.flags;
    int in_mode = IN_GLOBAL;
    if (f & SFH_LOAD_LDS) {
        load_via_lds<T, NT, E, MODE>(a, c, lds);
        in_mode = IN_LDS;
    }
    constexpr bool two = (MODE & SFH_M_MID) != 0;
    const uint32_t np = a.npass;
    const uint32_t total = two ? 2 * np - 1 : np; // FFT #2 pass 0 is fused into FFT #1's last
    for (uint32_t s = 0; s < total; ++s) {
        // fused program: FFT #1 passes 0..np-1, then FFT #2 passes 1..np-1 (stored at np+1..)
        const uint32_t pi = (s < np) ? s : s + 1;
        int out_mode;
        if (s + 1 == total) {
            out_mode = (f & SFH_STORE_LDS) ? OUT_LDS : OUT_GLOBAL;
        } else if (two && s + 1 == np) {
            out_mode = OUT_MID;
        } else {
            out_mode = OUT_LDS;
        }
        dispatch_pass<T, NT, E, MODE>(a, c, lds, a.pass[pi], s == 0 ? in_mode : IN_LDS, out_mode, wvalid);
    }
    if (f & SFH_STORE_LDS) {
        store_via_lds<T, NT, E, MODE>(a, c, lds);
    }
}

} // namespace sfh
