// Error text crossing a provider ABI: the caller passes a buffer of `capacity` bytes; the
// provider leaves a NUL-terminated message in it, cut to capacity - 1 characters (the
// convention of the reference's provider modules, include/syclfft/detail/provider_error_buffer.hpp).
#pragma once

#include <cstddef>
#include <string>

namespace syclfft::detail
{
    inline void set_error_buffer(char *output, std::size_t capacity, const std::string &message)
    {
        if (output != nullptr && capacity > 0) {
            output[message.copy(output, capacity - 1)] = '\0';
        }
    }
} // namespace syclfft::detail
