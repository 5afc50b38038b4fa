// A provider module whose ABI version mismatches is rejected with plugin_error
// (semantics of the reference's tests/plugin_tests.cpp).  Run with SYCLFFT_PLUGIN_PATH
// pointing at the directory holding the bad module.
#include <iostream>
#include <syclfft/host.hpp>

int main()
{
    try {
        (void)syclfft::host::plan_dft_1d<float>(4, syclfft::direction::forward);
    } catch (const syclfft::exception &ex) {
        if (ex.code() == syclfft::error_code::plugin_error) {
            std::cout << "plugin ABI test passed\n";
            return 0;
        }
        std::cerr << "unexpected error: " << ex.what() << '\n';
        return 1;
    }
    std::cerr << "provider with incompatible ABI was accepted\n";
    return 1;
}
