// A host provider module with an incompatible ABI version (the fake the reference tests
// with, tests/bad_fftw_provider.cpp): loading it must raise plugin_error.
#include <syclfft/detail/provider_abi.h>

static const syclfft_host_provider_v1 k_bad{SYCLFFT_PROVIDER_ABI_VERSION + 1, "bad-abi", nullptr, nullptr, nullptr};

SYCLFFT_PROVIDER_EXPORT const syclfft_host_provider_v1 *syclfft_get_host_provider_v1(void)
{
    return &k_bad;
}
