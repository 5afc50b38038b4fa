// Core: validation helpers and provider-module loading.
// Behaviour follows the reference (src/core/provider_loader.cpp:112-232): element-count
// checks, normalization scale, provider names, search order SYCLFFT_PLUGIN_PATH then
// <dir of this library>/syclfft/providers, dlopen(RTLD_NOW | RTLD_LOCAL), ABI check.
#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <limits>
#include <sstream>
#include <syclfft/detail/provider_loader.hpp>
#include <syclfft/detail/trace.hpp>

namespace syclfft::detail
{
    namespace
    {
        std::vector<std::string> split_paths(const char *value)
        {
            std::vector<std::string> out;
            if (!value) {
                return out;
            }
            std::stringstream in(value);
            std::string item;
            while (std::getline(in, item, ':')) {
                if (!item.empty()) {
                    out.push_back(item);
                }
            }
            return out;
        }

        std::filesystem::path this_library_dir()
        {
            Dl_info info{};
            if (dladdr(reinterpret_cast<void *>(&this_library_dir), &info) == 0 || !info.dli_fname) {
                return {};
            }
            return std::filesystem::path(info.dli_fname).parent_path();
        }
    } // namespace

    std::size_t checked_element_count(const std::vector<std::size_t> &lengths, std::size_t batch_count)
    {
        if (lengths.empty() || lengths.size() > 3) {
            throw exception(error_code::invalid_argument, "FFT rank must be between 1 and 3");
        }
        if (batch_count == 0) {
            throw exception(error_code::invalid_argument, "FFT batch count must be non-zero");
        }
        std::size_t count = batch_count;
        for (const auto n : lengths) {
            if (n == 0) {
                throw exception(error_code::invalid_argument, "FFT dimensions must be non-zero");
            }
            if (count > std::numeric_limits<std::size_t>::max() / n) {
                throw exception(error_code::invalid_argument, "FFT element count overflows size_t");
            }
            count *= n;
        }
        return count;
    }

    double normalization_scale(normalization mode, direction dir, std::size_t transform_size)
    {
        const double n = static_cast<double>(transform_size);
        switch (mode) {
        case normalization::none:
            return 1.0;
        case normalization::forward:
            return dir == direction::forward ? 1.0 / n : 1.0;
        case normalization::backward:
            return dir == direction::backward ? 1.0 / n : 1.0;
        case normalization::orthogonal:
            return 1.0 / std::sqrt(n);
        }
        return 1.0;
    }

    const char *provider_name(provider value) noexcept
    {
        switch (value) {
        case provider::automatic:
            return "automatic";
        case provider::portable_sycl:
            return "portable_sycl";
        case provider::fftw:
            return "fftw";
        case provider::cufft:
            return "cufft";
        case provider::rocfft:
            return "rocfft";
        case provider::onemkl:
            return "onemkl";
        case provider::hip_native:
            return "hip_native";
        }
        return "unknown";
    }

    std::vector<std::string> provider_search_paths()
    {
        auto out = split_paths(std::getenv("SYCLFFT_PLUGIN_PATH"));
        const auto dir = this_library_dir();
        if (!dir.empty()) {
            out.push_back((dir / "syclfft" / "providers").string());
        }
        return out;
    }

    loaded_provider_symbol load_provider_symbol(const std::string &provider_file, const std::string &symbol_name)
    {
        std::string errors;
        for (const auto &dir : provider_search_paths()) {
            const auto path = std::filesystem::path(dir) / (provider_file + ".so");
            void *raw = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (!raw) {
                const char *msg = dlerror();
                errors += path.string() + ": " + (msg ? msg : "load failed") + "; ";
                continue;
            }
            std::shared_ptr<void> module(raw, [](void *h) {
                if (h) {
                    dlclose(h);
                }
            });
            void *symbol = dlsym(raw, symbol_name.c_str());
            if (!symbol) {
                errors += path.string() + ": entry point '" + symbol_name + "' missing; ";
                continue;
            }
            trace("core: provider symbol resolved");
            return {symbol, std::move(module)};
        }
        throw exception(
            error_code::provider_unavailable, "Unable to load provider '" + provider_file + "': " + errors);
    }

    loaded_host_provider load_host_provider(const std::string &provider_file)
    {
        trace("core: resolving host provider symbol");
        auto loaded = load_provider_symbol(provider_file, "syclfft_get_host_provider_v1");
        const auto getter = reinterpret_cast<syclfft_get_host_provider_v1_fn>(loaded.symbol);
        const auto *api = getter();
        if (!api || api->abi_version != SYCLFFT_PROVIDER_ABI_VERSION) {
            throw exception(error_code::plugin_error, "Provider '" + provider_file + "' has an incompatible ABI");
        }
        return {api, std::move(loaded.module)};
    }

} // namespace syclfft::detail
