// Diagnostic trace: with SYCLFFT_TRACE set to a value other than "" or "0", each message
// goes to stderr as one line "[syclfft] <message>" (the reference's trace switch and
// prefix, include/syclfft/detail/trace.hpp:8-18).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace syclfft::detail
{
    inline bool trace_enabled() noexcept
    {
        const char *v = std::getenv("SYCLFFT_TRACE");
        return v != nullptr && std::strcmp(v, "") != 0 && std::strcmp(v, "0") != 0;
    }

    inline void trace(const char *message) noexcept
    {
        if (trace_enabled()) {
            // one write per line, so lines of concurrent callers do not interleave
            std::fprintf(stderr, "[syclfft] %s\n", message ? message : "");
            std::fflush(stderr);
        }
    }
} // namespace syclfft::detail
