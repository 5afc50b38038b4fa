// `[syclfft] <message>` on stderr when SYCLFFT_TRACE is set and not "0"
// (reference: include/syclfft/detail/trace.hpp:8-18).
#pragma once

#include <cstdio>
#include <cstdlib>

namespace syclfft::detail
{
    inline void trace(const char *message) noexcept
    {
        const char *enabled = std::getenv("SYCLFFT_TRACE");
        if (!enabled || enabled[0] == '\0' || enabled[0] == '0') {
            return;
        }
        std::fprintf(stderr, "[syclfft] %s\n", message);
        std::fflush(stderr);
    }
} // namespace syclfft::detail
