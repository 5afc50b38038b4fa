// Common types of the syclfft plan API, with the reference's names, values and meaning
// (reference: include/syclfft/common.hpp:34-128).  `provider::hip_native` is appended for
// the MI355X provider; every existing enumerator keeps its value.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <syclfft/span.hpp>
#include <vector>

#define SYCLFFT_EXPORT __attribute__((visibility("default")))

namespace syclfft
{

    enum class direction
    {
        forward = -1,
        backward = 1
    };

    enum class placement
    {
        in_place,
        out_of_place
    };

    enum class normalization
    {
        none,
        forward,
        backward,
        orthogonal
    };

    enum class planning_mode
    {
        estimate,
        measure
    };

    enum class provider
    {
        automatic,
        portable_sycl,
        fftw,
        cufft,
        rocfft,
        onemkl,
        hip_native
    };

    struct plan_options
    {
        syclfft::placement placement = syclfft::placement::out_of_place;
        syclfft::normalization normalization = syclfft::normalization::none;
        planning_mode planning = planning_mode::estimate;
        provider preferred_provider = provider::automatic;
    };

    struct provider_status
    {
        provider id{};
        std::string name;
        bool built{};
        bool available{};
        std::string reason;
    };

    enum class error_code
    {
        invalid_argument,
        invalid_state,
        invalid_pointer,
        unsupported_precision,
        provider_unavailable,
        planning_failed,
        execution_failed,
        plugin_error,
    };

    class SYCLFFT_EXPORT exception : public std::runtime_error
    {
    public:
        exception(error_code code, const std::string &message)
          : std::runtime_error(message)
          , code_(code)
        {}

        [[nodiscard]] error_code code() const noexcept
        {
            return code_;
        }

    private:
        error_code code_;
    };

    namespace detail
    {
        // rank 1..3, non-zero lengths and batch, no size_t overflow
        SYCLFFT_EXPORT std::size_t
            checked_element_count(const std::vector<std::size_t> &lengths, std::size_t batch_count);
        // 1, 1/N or 1/sqrt(N) by mode and direction; N = product of lengths
        SYCLFFT_EXPORT double
            normalization_scale(normalization mode, direction dir, std::size_t transform_size);
        SYCLFFT_EXPORT const char *provider_name(provider value) noexcept;

        inline std::size_t transform_size(syclfft::span<const std::size_t> shape)
        {
            std::size_t result = 1;
            for (const auto length : shape) {
                result *= length;
            }
            return result;
        }
    } // namespace detail

} // namespace syclfft
