/*
 * TEST INFRASTRUCTURE ONLY.  The CPU oracle of the MI355X syclfft provider.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker / the timed CPU baseline; the provider never does.
 *
 * Two families of functions:
 *
 *  1. FP64-quality oracle (long double arithmetic, exact integer exponent reduction):
 *     - oracle_dft_ld       brute-force O(n^2) DFT, X[k] = sum_j x[j] exp(sign 2 pi i jk/n),
 *                           exponent jk reduced mod n in integers.  Pinned by the known
 *                           answers in tests/test_oracle.py (impulse, shifted impulse, tone,
 *                           constant) and by scipy/numpy fixtures.
 *     - oracle_fft_ld       fast transform of any length: recursive mixed radix for
 *                           7-smooth n, Bluestein with a power-of-two convolution otherwise.
 *                           Pinned against oracle_dft_ld and the golden fixtures.
 *     - oracle_fftn_ld      rank 1..3 batched C-order transform, per axis (the axis loop of
 *                           src/sycl/sycl.cpp:692-767), plus normalization_scale
 *                           (src/core/provider_loader.cpp:134-147).
 *
 *  2. Restatement of the reference's portable provider (src/sycl/sycl.cpp), in the plan's
 *     Scalar precision with the reference's own tables and kernel sequence, so its
 *     rounding (Scalar twiddles, unreduced chirp exponent) is reproduced:
 *     - ref_portable_exec_f32 / _f64: plan tables (sycl.cpp:509-569), per-axis radix-2
 *       (sycl.cpp:175-234), mixed-radix (sycl.cpp:100-173) or Bluestein (sycl.cpp:268-339)
 *       pass sequence, then copy_scaled (sycl.cpp:87-98, 769-773).  Parallel loops use
 *       OpenMP the way AdaptiveCpp's OpenMP backend runs each parallel_for.
 *     Parity of this restatement: tests/test_oracle.py checks it against the FP64 oracle
 *     on every path (pow2, 7-smooth, Bluestein, multi-D) at the reference's tolerances
 *     (tests/sycl_tests.cpp:186, 242); its Bluestein fp32 error is the one the reference's
 *     Scalar-precision chirp produces.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct
{
    long double re, im;
} cl;

static const long double k_two_pi = 6.283185307179586476925286766559005768L;

/* exp(sign * 2 pi i m / n) with m reduced mod n */
static cl root(uint64_t m, uint64_t n, int sign)
{
    m %= n;
    long double a = k_two_pi * (long double)m / (long double)n;
    cl r = {cosl(a), sinl(a) * (long double)sign};
    return r;
}

static inline cl cmul(cl a, cl b)
{
    cl r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* ---------------------------------------------------------------- brute force */

void oracle_dft_ld(uint64_t n, const double *in, double *out, int sign)
{
    for (uint64_t k = 0; k < n; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (uint64_t j = 0; j < n; ++j) {
            const unsigned __int128 e = (unsigned __int128)j * k;
            const cl w = root((uint64_t)(e % n), n, sign);
            const long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * w.re - xi * w.im;
            si += xr * w.im + xi * w.re;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ---------------------------------------------------------------- fast long double */

static int smallest_factor7(uint64_t n)
{
    static const int f[] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; ++i) {
        if (n % (uint64_t)f[i] == 0) {
            return f[i];
        }
    }
    return 0;
}

static int is_smooth7(uint64_t n)
{
    static const uint64_t f[] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i) {
        while (n % f[i] == 0) {
            n /= f[i];
        }
    }
    return n == 1;
}

/* decimation in time: out[0..n) = DFT of x[0], x[s], x[2s], ... using roots of the top length */
static void fft_rec(const cl *x, uint64_t s, uint64_t n, cl *out, const cl *roots, uint64_t top,
                    uint64_t step, cl *tmp)
{
    if (n == 1) {
        out[0] = x[0];
        return;
    }
    const int r = smallest_factor7(n);
    const uint64_t m = n / (uint64_t)r;
    for (int q = 0; q < r; ++q) {
        fft_rec(x + (uint64_t)q * s, s * (uint64_t)r, m, out + (uint64_t)q * m, roots, top,
                step * (uint64_t)r, tmp);
    }
    for (uint64_t k = 0; k < m; ++k) {
        cl v[7];
        for (int q = 0; q < r; ++q) {
            v[q] = cmul(out[(uint64_t)q * m + k], roots[(step * (uint64_t)q * k) % top]);
        }
        for (int p = 0; p < r; ++p) {
            cl acc = {0.0L, 0.0L};
            for (int q = 0; q < r; ++q) {
                const cl w = roots[(step * m * (uint64_t)((q * p) % r)) % top];
                const cl t = cmul(v[q], w);
                acc.re += t.re;
                acc.im += t.im;
            }
            tmp[k + (uint64_t)p * m] = acc;
        }
    }
    memcpy(out, tmp, n * sizeof(cl));
}

static void fft_smooth(cl *data, uint64_t n, int sign)
{
    cl *roots = (cl *)malloc(n * sizeof(cl));
    cl *out = (cl *)malloc(n * sizeof(cl));
    cl *tmp = (cl *)malloc(n * sizeof(cl));
    for (uint64_t m = 0; m < n; ++m) {
        roots[m] = root(m, n, sign);
    }
    fft_rec(data, 1, n, out, roots, n, 1, tmp);
    memcpy(data, out, n * sizeof(cl));
    free(roots);
    free(out);
    free(tmp);
}

static void fft_any(cl *data, uint64_t n, int sign)
{
    if (n <= 1) {
        return;
    }
    if (is_smooth7(n)) {
        fft_smooth(data, n, sign);
        return;
    }
    /* Bluestein, power-of-two convolution length (independent of the provider's choice) */
    uint64_t M = 1;
    while (M < 2 * n - 1) {
        M *= 2;
    }
    cl *w = (cl *)malloc(n * sizeof(cl));
    cl *a = (cl *)calloc(M, sizeof(cl));
    cl *b = (cl *)calloc(M, sizeof(cl));
    for (uint64_t j = 0; j < n; ++j) {
        /* w[j] = exp(sign i pi j^2 / n), j^2 reduced mod 2n */
        const unsigned __int128 sq = (unsigned __int128)j * j;
        w[j] = root((uint64_t)(sq % (2 * (unsigned __int128)n)), 2 * n, sign);
        a[j] = cmul(data[j], w[j]);
        const cl cw = {w[j].re, -w[j].im};
        b[j] = cw;
        if (j) {
            b[M - j] = cw;
        }
    }
    fft_smooth(a, M, -1);
    fft_smooth(b, M, -1);
    for (uint64_t i = 0; i < M; ++i) {
        a[i] = cmul(a[i], b[i]);
    }
    fft_smooth(a, M, 1);
    const long double inv = 1.0L / (long double)M;
    for (uint64_t k = 0; k < n; ++k) {
        cl v = cmul(a[k], w[k]);
        data[k].re = v.re * inv;
        data[k].im = v.im * inv;
    }
    free(w);
    free(a);
    free(b);
}

void oracle_fft_ld(uint64_t n, const double *in, double *out, int sign)
{
    cl *d = (cl *)malloc(n * sizeof(cl));
    for (uint64_t i = 0; i < n; ++i) {
        d[i].re = in[2 * i];
        d[i].im = in[2 * i + 1];
    }
    fft_any(d, n, sign);
    for (uint64_t i = 0; i < n; ++i) {
        out[2 * i] = (double)d[i].re;
        out[2 * i + 1] = (double)d[i].im;
    }
    free(d);
}

double oracle_norm_scale(int norm, int sign, uint64_t n)
{
    switch (norm) {
    case 0:
        return 1.0;
    case 1:
        return sign == -1 ? 1.0 / (double)n : 1.0;
    case 2:
        return sign == 1 ? 1.0 / (double)n : 1.0;
    case 3:
        return 1.0 / sqrt((double)n);
    default:
        return 1.0;
    }
}

/* rank 1..3, C order, batch outermost; in/out interleaved double (may alias) */
void oracle_fftn_ld(int rank, const uint64_t *lengths, uint64_t batch, const double *in,
                    double *out, int sign, int norm)
{
    uint64_t tsize = 1;
    for (int a = 0; a < rank; ++a) {
        tsize *= lengths[a];
    }
    const uint64_t total = tsize * batch;
    cl *d = (cl *)malloc(total * sizeof(cl));
    for (uint64_t i = 0; i < total; ++i) {
        d[i].re = in[2 * i];
        d[i].im = in[2 * i + 1];
    }
    for (int a = 0; a < rank; ++a) {
        const uint64_t n = lengths[a];
        uint64_t inner = 1;
        for (int k = a + 1; k < rank; ++k) {
            inner *= lengths[k];
        }
        const uint64_t outer = total / (n * inner);
#pragma omp parallel
        {
            cl *line = (cl *)malloc(n * sizeof(cl));
#pragma omp for schedule(dynamic, 1)
            for (int64_t l = 0; l < (int64_t)(outer * inner); ++l) {
                const uint64_t o = (uint64_t)l / inner, i = (uint64_t)l % inner;
                const uint64_t base = o * n * inner + i;
                for (uint64_t j = 0; j < n; ++j) {
                    line[j] = d[base + j * inner];
                }
                fft_any(line, n, sign);
                for (uint64_t j = 0; j < n; ++j) {
                    d[base + j * inner] = line[j];
                }
            }
            free(line);
        }
    }
    const long double scale = (long double)oracle_norm_scale(norm, sign, tsize);
    for (uint64_t i = 0; i < total; ++i) {
        out[2 * i] = (double)(d[i].re * scale);
        out[2 * i + 1] = (double)(d[i].im * scale);
    }
    free(d);
}

/* ---------------------------------------------------------------- reference restatement */

#define REF_DEFINE(SUF, S, COS, SIN, PI_S)                                                         \
    typedef struct                                                                                 \
    {                                                                                              \
        S re, im;                                                                                  \
    } rc_##SUF;                                                                                    \
    static inline rc_##SUF rmul_##SUF(rc_##SUF a, rc_##SUF b)                                      \
    {                                                                                              \
        rc_##SUF r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};                       \
        return r;                                                                                  \
    }                                                                                              \
    /* make_twiddles, src/sycl/sycl.cpp:509-518 (Scalar precision) */                              \
    static rc_##SUF *ref_twiddles_##SUF(uint64_t len, S sign)                                      \
    {                                                                                              \
        rc_##SUF *t = (rc_##SUF *)malloc(len * sizeof(rc_##SUF));                                  \
        const S two_pi = (S)2 * PI_S;                                                              \
        for (uint64_t i = 0; i < len; ++i) {                                                       \
            const S ang = sign * two_pi * (S)i / (S)len;                                           \
            t[i].re = COS(ang);                                                                    \
            t[i].im = SIN(ang);                                                                    \
        }                                                                                          \
        return t;                                                                                  \
    }                                                                                              \
    /* enqueue_radix2_axis, src/sycl/sycl.cpp:175-234 */                                           \
    static void ref_radix2_##SUF(const rc_##SUF *in, rc_##SUF *out, uint64_t total, uint64_t len, \
                                 uint64_t stride, const rc_##SUF *tw, int conj)                    \
    {                                                                                              \
        unsigned bits = 0;                                                                         \
        for (uint64_t v = len; v > 1; v >>= 1) {                                                   \
            ++bits;                                                                                \
        }                                                                                          \
        _Pragma("omp parallel for") for (int64_t li = 0; li < (int64_t)total; ++li)                \
        {                                                                                          \
            const uint64_t lin = (uint64_t)li;                                                     \
            const uint64_t c = (lin / stride) % len;                                               \
            uint64_t rev = 0, v = c;                                                               \
            for (unsigned b = 0; b < bits; ++b) {                                                  \
                rev = (rev << 1) | (v & 1);                                                        \
                v >>= 1;                                                                           \
            }                                                                                      \
            out[lin] = in[lin - c * stride + rev * stride];                                        \
        }                                                                                          \
        const uint64_t lines = total / len;                                                        \
        for (uint64_t stage = 2; stage <= len; stage <<= 1) {                                      \
            _Pragma("omp parallel for") for (int64_t pi = 0; pi < (int64_t)(lines * (len / 2));    \
                                             ++pi)                                                 \
            {                                                                                      \
                const uint64_t pair = (uint64_t)pi;                                                \
                const uint64_t line = pair / (len / 2), pil = pair % (len / 2);                    \
                const uint64_t half = stage / 2, group = pil / half, j = pil % half;               \
                const uint64_t base = (line / stride) * len * stride + line % stride;              \
                const uint64_t first = base + (group * stage + j) * stride;                        \
                const uint64_t second = first + half * stride;                                     \
                rc_##SUF t = tw[j * (len / stage)];                                                \
                if (conj) {                                                                        \
                    t.im = -t.im;                                                                  \
                }                                                                                  \
                const rc_##SUF lhs = out[first];                                                   \
                const rc_##SUF rhs = rmul_##SUF(out[second], t);                                   \
                out[first].re = lhs.re + rhs.re;                                                   \
                out[first].im = lhs.im + rhs.im;                                                   \
                out[second].re = lhs.re - rhs.re;                                                  \
                out[second].im = lhs.im - rhs.im;                                                  \
            }                                                                                      \
        }                                                                                          \
    }                                                                                              \
    /* enqueue_mixed_axis, src/sycl/sycl.cpp:100-173 */                                            \
    static void ref_mixed_##SUF(const rc_##SUF *in, rc_##SUF *out, uint64_t total, uint64_t len,  \
                                uint64_t stride, const rc_##SUF *tw, const uint64_t *fac, int nf)  \
    {                                                                                              \
        _Pragma("omp parallel for") for (int64_t li = 0; li < (int64_t)total; ++li)                \
        {                                                                                          \
            const uint64_t lin = (uint64_t)li;                                                     \
            const uint64_t c = (lin / stride) % len;                                               \
            const uint64_t base = lin - c * stride;                                                \
            uint64_t src = 0, rem = len, dig = c;                                                  \
            for (int s = 0; s < nf; ++s) {                                                         \
                rem /= fac[s];                                                                     \
                src += (dig % fac[s]) * rem;                                                       \
                dig /= fac[s];                                                                     \
            }                                                                                      \
            out[lin] = in[base + src * stride];                                                    \
        }                                                                                          \
        const uint64_t lines = total / len;                                                        \
        uint64_t ssize = 1;                                                                        \
        for (int s = 0; s < nf; ++s) {                                                             \
            const uint64_t prev = ssize, r = fac[s];                                               \
            ssize *= r;                                                                            \
            const uint64_t st = ssize;                                                             \
            _Pragma("omp parallel for") for (int64_t bi = 0; bi < (int64_t)(lines * (len / r));    \
                                             ++bi)                                                 \
            {                                                                                      \
                const uint64_t bf = (uint64_t)bi;                                                  \
                const uint64_t line = bf / (len / r), inl = bf % (len / r);                        \
                const uint64_t group = inl / prev, j = inl % prev;                                 \
                const uint64_t base = (line / stride) * len * stride + line % stride               \
                                      + group * st * stride + j * stride;                          \
                rc_##SUF v[7];                                                                     \
                for (uint64_t q = 0; q < r; ++q) {                                                 \
                    v[q] = rmul_##SUF(out[base + q * prev * stride], tw[(q * j * (len / st)) % len]); \
                }                                                                                  \
                for (uint64_t p = 0; p < r; ++p) {                                                 \
                    rc_##SUF sum = {(S)0, (S)0};                                                   \
                    for (uint64_t q = 0; q < r; ++q) {                                             \
                        const rc_##SUF t = rmul_##SUF(v[q], tw[(q * p * (len / r)) % len]);        \
                        sum.re += t.re;                                                            \
                        sum.im += t.im;                                                            \
                    }                                                                              \
                    out[base + p * prev * stride] = sum;                                           \
                }                                                                                  \
            }                                                                                      \
        }                                                                                          \
    }                                                                                              \
    /* host_radix2_fft, src/sycl/sycl.cpp:236-266 (forward, per-butterfly cos/sin) */             \
    static void ref_host_radix2_##SUF(rc_##SUF *v, uint64_t len)                                   \
    {                                                                                              \
        for (uint64_t i = 1, rev = 0; i < len; ++i) {                                              \
            uint64_t bit = len >> 1;                                                               \
            while (rev & bit) {                                                                    \
                rev ^= bit;                                                                        \
                bit >>= 1;                                                                         \
            }                                                                                      \
            rev ^= bit;                                                                            \
            if (i < rev) {                                                                         \
                rc_##SUF t = v[i];                                                                 \
                v[i] = v[rev];                                                                     \
                v[rev] = t;                                                                        \
            }                                                                                      \
        }                                                                                          \
        const S two_pi = (S)2 * PI_S;                                                              \
        for (uint64_t stage = 2; stage <= len; stage *= 2) {                                       \
            for (uint64_t b = 0; b < len; b += stage) {                                            \
                for (uint64_t j = 0; j < stage / 2; ++j) {                                         \
                    const S ang = (S)(-1) * two_pi * (S)j / (S)stage;                              \
                    const rc_##SUF t = {COS(ang), SIN(ang)};                                       \
                    const rc_##SUF lhs = v[b + j];                                                 \
                    const rc_##SUF rhs = rmul_##SUF(v[b + j + stage / 2], t);                      \
                    v[b + j].re = lhs.re + rhs.re;                                                 \
                    v[b + j].im = lhs.im + rhs.im;                                                 \
                    v[b + j + stage / 2].re = lhs.re - rhs.re;                                     \
                    v[b + j + stage / 2].im = lhs.im - rhs.im;                                     \
                }                                                                                  \
            }                                                                                      \
        }                                                                                          \
    }                                                                                              \
    static int ref_factors_##SUF(uint64_t n, uint64_t *fac)                                        \
    {                                                                                              \
        static const uint64_t rs[] = {2, 3, 5, 7};                                                 \
        int nf = 0;                                                                                \
        for (int i = 0; i < 4; ++i) {                                                              \
            while (n % rs[i] == 0) {                                                               \
                fac[nf++] = rs[i];                                                                 \
                n /= rs[i];                                                                        \
            }                                                                                      \
        }                                                                                          \
        return n == 1 ? nf : -1;                                                                   \
    }                                                                                              \
    /* plan<Scalar>::impl::execute portable branch, src/sycl/sycl.cpp:688-775 */                   \
    int ref_portable_exec_##SUF(int rank, const uint64_t *lengths, uint64_t batch, int dir,        \
                                int norm, const S *in_s, S *out_s)                                 \
    {                                                                                              \
        uint64_t tsize = 1;                                                                        \
        for (int a = 0; a < rank; ++a) {                                                           \
            tsize *= lengths[a];                                                                   \
        }                                                                                          \
        const uint64_t total = tsize * batch;                                                      \
        const S sign = dir == -1 ? (S)(-1) : (S)1;                                                 \
        uint64_t scratch = total;                                                                  \
        for (int a = 0; a < rank; ++a) {                                                           \
            uint64_t fac[64];                                                                      \
            if (lengths[a] != 1 && ref_factors_##SUF(lengths[a], fac) < 0) {                       \
                uint64_t M = 1;                                                                    \
                while (M < 2 * lengths[a] - 1) {                                                   \
                    M *= 2;                                                                        \
                }                                                                                  \
                if (total / lengths[a] * M > scratch) {                                            \
                    scratch = total / lengths[a] * M;                                              \
                }                                                                                  \
            }                                                                                      \
        }                                                                                          \
        rc_##SUF *A = (rc_##SUF *)malloc(scratch * sizeof(rc_##SUF));                              \
        rc_##SUF *B = (rc_##SUF *)malloc(scratch * sizeof(rc_##SUF));                              \
        if (!A || !B) {                                                                            \
            free(A);                                                                               \
            free(B);                                                                               \
            return -1;                                                                             \
        }                                                                                          \
        const rc_##SUF *cur = (const rc_##SUF *)in_s;                                              \
        for (int a = 0; a < rank; ++a) {                                                           \
            const uint64_t len = lengths[a];                                                       \
            uint64_t stride = 1;                                                                   \
            for (int k = a + 1; k < rank; ++k) {                                                   \
                stride *= lengths[k];                                                              \
            }                                                                                      \
            uint64_t fac[64];                                                                      \
            const int nf = ref_factors_##SUF(len, fac);                                            \
            if (nf >= 0) {                                                                         \
                rc_##SUF *tw = ref_twiddles_##SUF(len, sign);                                      \
                rc_##SUF *dst = (cur == A) ? B : A;                                                \
                if (len == 1) {                                                                    \
                    memmove(dst, cur, total * sizeof(rc_##SUF));                                   \
                } else if ((len & (len - 1)) == 0) {                                               \
                    ref_radix2_##SUF(cur, dst, total, len, stride, tw, 0);                         \
                } else {                                                                           \
                    ref_mixed_##SUF(cur, dst, total, len, stride, tw, fac, nf);                    \
                }                                                                                  \
                free(tw);                                                                          \
                cur = dst;                                                                         \
            } else {                                                                               \
                /* Bluestein tables, src/sycl/sycl.cpp:530-568 */                                 \
                uint64_t M = 1;                                                                    \
                while (M < 2 * len - 1) {                                                          \
                    M *= 2;                                                                        \
                }                                                                                  \
                rc_##SUF *kern = (rc_##SUF *)calloc(M, sizeof(rc_##SUF));                          \
                rc_##SUF *chirp = (rc_##SUF *)malloc(len * sizeof(rc_##SUF));                      \
                for (uint64_t i = 0; i < len; ++i) {                                               \
                    const S sq = (S)i * (S)i;                                                      \
                    const S ang = -sign * PI_S * sq / (S)len;                                      \
                    const rc_##SUF v = {COS(ang), SIN(ang)};                                       \
                    kern[i] = v;                                                                   \
                    if (i) {                                                                       \
                        kern[M - i] = v;                                                           \
                    }                                                                              \
                    chirp[i].re = v.re;                                                            \
                    chirp[i].im = -v.im;                                                           \
                }                                                                                  \
                ref_host_radix2_##SUF(kern, M);                                                    \
                rc_##SUF *ctw = ref_twiddles_##SUF(M, (S)(-1));                                    \
                rc_##SUF *first = (cur == A) ? B : A;                                              \
                rc_##SUF *second = (first == A) ? B : A;                                           \
                const uint64_t lines = total / len, exp = lines * M;                               \
                /* expand, sycl.cpp:279-294 */                                                     \
                _Pragma("omp parallel for") for (int64_t ei = 0; ei < (int64_t)exp; ++ei)          \
                {                                                                                  \
                    const uint64_t e = (uint64_t)ei, line = e / M, c = e % M;                      \
                    if (c < len) {                                                                 \
                        const uint64_t srcix =                                                     \
                            (line / stride) * len * stride + line % stride + c * stride;           \
                        first[e] = rmul_##SUF(cur[srcix], chirp[c]);                               \
                    } else {                                                                       \
                        first[e].re = (S)0;                                                        \
                        first[e].im = (S)0;                                                        \
                    }                                                                              \
                }                                                                                  \
                ref_radix2_##SUF(first, second, exp, M, 1, ctw, 0);                                \
                _Pragma("omp parallel for") for (int64_t ei = 0; ei < (int64_t)exp; ++ei)          \
                {                                                                                  \
                    second[ei] = rmul_##SUF(second[ei], kern[(uint64_t)ei % M]);                   \
                }                                                                                  \
                ref_radix2_##SUF(second, first, exp, M, 1, ctw, 1);                                \
                const S inv = (S)1 / (S)M;                                                         \
                _Pragma("omp parallel for") for (int64_t li = 0; li < (int64_t)total; ++li)        \
                {                                                                                  \
                    const uint64_t lin = (uint64_t)li, c = (lin / stride) % len;                   \
                    const uint64_t line = (lin / (len * stride)) * stride + lin % stride;          \
                    const rc_##SUF ch = {chirp[c].re * inv, chirp[c].im * inv};                    \
                    second[lin] = rmul_##SUF(first[line * M + c], ch);                             \
                }                                                                                  \
                free(kern);                                                                        \
                free(chirp);                                                                       \
                free(ctw);                                                                         \
                cur = second;                                                                      \
            }                                                                                      \
        }                                                                                          \
        /* copy_scaled, sycl.cpp:87-98, 769-773 */                                                 \
        const S scale = (S)oracle_norm_scale(norm, dir, tsize);                                    \
        rc_##SUF *out = (rc_##SUF *)out_s;                                                         \
        _Pragma("omp parallel for") for (int64_t i = 0; i < (int64_t)total; ++i)                   \
        {                                                                                          \
            out[i].re = cur[i].re * scale;                                                         \
            out[i].im = cur[i].im * scale;                                                         \
        }                                                                                          \
        free(A);                                                                                   \
        free(B);                                                                                   \
        return 0;                                                                                  \
    }

REF_DEFINE(f32, float, cosf, sinf, 3.14159265358979323846f)
REF_DEFINE(f64, double, cos, sin, 3.14159265358979323846)

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) {
        omp_set_num_threads(n);
    }
#else
    (void)n;
#endif
}
