// Plan-time tables, generated in long double with exact integer index reduction.
//
// The reference builds its twiddles and Bluestein chirps in the plan's own Scalar
// precision, with the chirp exponent k^2 formed in Scalar and never reduced mod 2N
// (src/sycl/sycl.cpp:509-518, 546-560), and transforms the Bluestein kernel with a
// Scalar-precision radix-2 FFT whose twiddles come from per-butterfly cos/sin
// (src/sycl/sycl.cpp:236-266).  Here every table entry is computed from an exactly
// reduced integer exponent in 80-bit long double and rounded once to the plan's type.
#include "fft_tables.h"

#include <cmath>
#include <stdexcept>

namespace sfh
{

namespace
{
const long double k_two_pi = 6.283185307179586476925286766559005768L;
const long double k_pi = 3.141592653589793238462643383279502884L;
} // namespace

cxl root_ld(uint64_t m, uint64_t n)
{
    m %= n;
    // exact symmetry points keep the table exact where the value is exactly representable
    if (m == 0) {
        return {1.0L, 0.0L};
    }
    if (2 * m == n) {
        return {-1.0L, 0.0L};
    }
    if (4 * m == n) {
        return {0.0L, -1.0L};
    }
    if (4 * m == 3 * n) {
        return {0.0L, 1.0L};
    }
    const long double angle = k_two_pi * static_cast<long double>(m) / static_cast<long double>(n);
    return {cosl(angle), -sinl(angle)};
}

cxl chirp_ld(uint64_t j, uint64_t n)
{
    // exp(-i pi j^2 / n) with j^2 reduced mod 2n in exact integer arithmetic
    const unsigned __int128 sq = static_cast<unsigned __int128>(j) * j;
    const uint64_t r = static_cast<uint64_t>(sq % (2 * static_cast<unsigned __int128>(n)));
    // exp(-i pi r / n) = root(r, 2n)
    return root_ld(r, 2 * n);
}

namespace
{
// Recursive mixed-radix (2,3,5,7) decimation-in-time forward DFT in long double.
void dft_rec(const cxl *x, size_t stride, size_t n, cxl *out, const std::vector<cxl> &roots,
             size_t root_step)
{
    if (n == 1) {
        out[0] = x[0];
        return;
    }
    size_t r = 0;
    for (size_t cand : {4u, 2u, 3u, 5u, 7u}) {
        if (n % cand == 0) {
            r = cand;
            break;
        }
    }
    if (r == 0) {
        throw std::runtime_error("host DFT length is not 7-smooth");
    }
    const size_t m = n / r;
    // sub-transforms of x[q + r*s] into out[q*m .. q*m+m)
    for (size_t q = 0; q < r; ++q) {
        dft_rec(x + q * stride, stride * r, m, out + q * m, roots, root_step * r);
    }
    std::vector<cxl> tmp(n);
    const size_t big_n = roots.size();
    for (size_t k = 0; k < m; ++k) {
        for (size_t p = 0; p < r; ++p) {
            const size_t kk = k + p * m;
            cxl acc{0.0L, 0.0L};
            for (size_t q = 0; q < r; ++q) {
                const cxl w = roots[(root_step * q * kk) % big_n];
                const cxl v = out[q * m + k];
                acc.re += v.re * w.re - v.im * w.im;
                acc.im += v.re * w.im + v.im * w.re;
            }
            tmp[kk] = acc;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        out[i] = tmp[i];
    }
}
} // namespace

std::vector<cxl> host_dft_ld(const std::vector<cxl> &x)
{
    const size_t n = x.size();
    std::vector<cxl> roots(n);
    for (size_t m = 0; m < n; ++m) {
        roots[m] = root_ld(m, n);
    }
    std::vector<cxl> out(n);
    dft_rec(x.data(), 1, n, out.data(), roots, 1);
    return out;
}

} // namespace sfh
