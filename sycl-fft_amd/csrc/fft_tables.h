// Plan-time table generation (long double, exact integer exponent reduction).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sfh
{

struct cxl
{
    long double re, im;
};

// exp(-2 pi i m / n), m reduced mod n exactly.
cxl root_ld(uint64_t m, uint64_t n);

// Bluestein chirp exp(-i pi j^2 / n) with j^2 reduced mod 2n exactly.
cxl chirp_ld(uint64_t j, uint64_t n);

// Forward DFT (exp(-2 pi i jk/n)) of a 7-smooth length, long double, plan time only.
std::vector<cxl> host_dft_ld(const std::vector<cxl> &x);

} // namespace sfh
