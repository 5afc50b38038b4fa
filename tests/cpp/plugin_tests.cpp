// Host ABI version check: with SYCLFFT_PLUGIN_PATH naming a directory whose
// syclfft_provider_fftw module reports another ABI version, planning in either precision
// fails with plugin_error, and the message names the module (reference behaviour:
// src/core/provider_loader.cpp:229-231, tests/plugin_tests.cpp).
#include <iostream>
#include <string>
#include <syclfft/host.hpp>

namespace
{
    template <class Scalar>
    bool refused(const char *label)
    {
        try {
            auto accepted = syclfft::host::plan_dft_1d<Scalar>(4, syclfft::direction::forward);
            (void)accepted;
        } catch (const syclfft::exception &ex) {
            const std::string text = ex.what();
            if (ex.code() != syclfft::error_code::plugin_error) {
                std::cerr << label << ": wrong error (" << static_cast<int>(ex.code()) << "): " << text << '\n';
                return false;
            }
            if (text.find("syclfft_provider_fftw") == std::string::npos) {
                std::cerr << label << ": message does not name the module: " << text << '\n';
                return false;
            }
            return true;
        }
        std::cerr << label << ": a module with another ABI version was accepted\n";
        return false;
    }
} // namespace

int main()
{
    const bool ok = refused<float>("float plan") && refused<double>("double plan");
    if (ok) {
        std::cout << "plugin ABI test passed\n";
    }
    return ok ? 0 : 1;
}
