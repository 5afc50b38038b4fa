// Scaffolding of the C++ plan-API test programs: checks that count themselves, plan
// options assembled by name, and an expectation on the error code a call raises.
#pragma once

#include <stdexcept>
#include <string>
#include <syclfft/common.hpp>

namespace tu
{
    inline int &checks_run()
    {
        static int n = 0;
        return n;
    }

    inline void check(bool ok, const std::string &what)
    {
        ++checks_run();
        if (!ok) {
            throw std::runtime_error("check failed: " + what);
        }
    }

    // plan_options by name, e.g. tu::opts().in_place().norm(normalization::forward).on(provider::fftw)
    class opts
    {
    public:
        opts &placed(syclfft::placement p)
        {
            o_.placement = p;
            return *this;
        }
        opts &in_place()
        {
            return placed(syclfft::placement::in_place);
        }
        opts &norm(syclfft::normalization n)
        {
            o_.normalization = n;
            return *this;
        }
        opts &on(syclfft::provider p)
        {
            o_.preferred_provider = p;
            return *this;
        }
        opts &measured()
        {
            o_.planning = syclfft::planning_mode::measure;
            return *this;
        }
        operator syclfft::plan_options() const
        {
            return o_;
        }

    private:
        syclfft::plan_options o_{};
    };

    // f() must throw syclfft::exception carrying `code`; returns the exception's message.
    template <class F>
    std::string expect_error(syclfft::error_code code, F &&f, const std::string &what)
    {
        std::string message;
        bool raised = false;
        try {
            f();
        } catch (const syclfft::exception &ex) {
            raised = true;
            message = ex.what();
            check(ex.code() == code, what + ": error code " + std::to_string(static_cast<int>(ex.code())) + " ("
                                         + message + "), expected " + std::to_string(static_cast<int>(code)));
        }
        check(raised, what + ": no syclfft::exception");
        return message;
    }
} // namespace tu
