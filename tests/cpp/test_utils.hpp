// Check helpers of the C++ API tests (same contract as the reference's tests/test_utils.hpp:9-40).
#pragma once

#include <stdexcept>
#include <string>
#include <syclfft/common.hpp>

namespace tu
{
    inline void check(bool condition, const std::string &message)
    {
        if (!condition) {
            throw std::runtime_error(message);
        }
    }

    inline syclfft::plan_options options(
        syclfft::placement placement = syclfft::placement::out_of_place,
        syclfft::normalization normalization = syclfft::normalization::none,
        syclfft::provider provider = syclfft::provider::automatic,
        syclfft::planning_mode planning = syclfft::planning_mode::estimate)
    {
        syclfft::plan_options o;
        o.placement = placement;
        o.normalization = normalization;
        o.preferred_provider = provider;
        o.planning = planning;
        return o;
    }

    template <class F>
    void expect_error(syclfft::error_code code, F &&f, const std::string &what)
    {
        try {
            f();
        } catch (const syclfft::exception &ex) {
            check(ex.code() == code, what + ": wrong error code (" + ex.what() + ")");
            return;
        }
        throw std::runtime_error(what + ": expected syclfft::exception");
    }
} // namespace tu
