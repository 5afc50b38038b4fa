// Core library: shape validation, normalization factors, provider names, and the loading
// of provider modules.
//
// Behaviour contract of the reference's core (src/core/provider_loader.cpp:112-232), which
// this file implements with its own structure:
//   * shapes: rank 1..3, every length and the batch count non-zero, the element count
//     representable in size_t, else invalid_argument;
//   * normalization: none -> 1; forward -> 1/N on forward plans; backward -> 1/N on backward
//     plans; orthogonal -> 1/sqrt(N) on both (N = product of the lengths);
//   * modules are `<name>.so` files, looked up in every SYCLFFT_PLUGIN_PATH entry
//     (':'-separated, empty entries skipped) and then in `syclfft/providers` next to this
//     library, never in the working directory; each candidate is opened with
//     dlopen(RTLD_NOW | RTLD_LOCAL) and must export the entry point;
//   * no usable candidate -> provider_unavailable (the message lists every attempt); a host
//     module whose table is missing or has another ABI version -> plugin_error.
#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <limits>
#include <string_view>
#include <syclfft/detail/provider_loader.hpp>
#include <syclfft/detail/trace.hpp>

namespace syclfft::detail
{
    namespace
    {
        // Directory holding this shared library ("" if the dynamic linker cannot say).
        std::string own_directory()
        {
            static const int anchor = 0;
            Dl_info where{};
            if (dladdr(&anchor, &where) == 0 || where.dli_fname == nullptr) {
                return {};
            }
            const std::string_view file(where.dli_fname);
            const auto slash = file.rfind('/');
            return slash == std::string_view::npos ? std::string(".") : std::string(file.substr(0, slash));
        }

        // An opened module; closing is the shared owner's job.
        std::shared_ptr<void> open_module(const std::string &path, std::string &why)
        {
            void *raw = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (raw == nullptr) {
                const char *msg = dlerror();
                why = msg ? msg : "dlopen failed";
                return {};
            }
            return std::shared_ptr<void>(raw, [](void *h) { dlclose(h); });
        }
    } // namespace

    std::size_t checked_element_count(const std::vector<std::size_t> &lengths, std::size_t batch_count)
    {
        const std::size_t rank = lengths.size();
        if (rank < 1 || rank > 3) {
            throw exception(error_code::invalid_argument, "a transform has 1, 2 or 3 dimensions");
        }
        if (batch_count == 0) {
            throw exception(error_code::invalid_argument, "the batch count must be at least 1");
        }
        constexpr std::size_t top = std::numeric_limits<std::size_t>::max();
        std::size_t total = batch_count;
        for (std::size_t axis = 0; axis < rank; ++axis) {
            const std::size_t n = lengths[axis];
            if (n == 0) {
                throw exception(error_code::invalid_argument, "length " + std::to_string(axis) + " is zero");
            }
            if (total > top / n) {
                throw exception(error_code::invalid_argument, "the element count does not fit in size_t");
            }
            total *= n;
        }
        return total;
    }

    double normalization_scale(normalization mode, direction dir, std::size_t transform_size)
    {
        const double n = static_cast<double>(transform_size);
        const bool divides = (mode == normalization::forward && dir == direction::forward)
                             || (mode == normalization::backward && dir == direction::backward);
        if (mode == normalization::orthogonal) {
            return 1.0 / std::sqrt(n);
        }
        return divides ? 1.0 / n : 1.0;
    }

    const char *provider_name(provider value) noexcept
    {
        static constexpr const char *names[] = {"automatic", "portable_sycl", "fftw",      "cufft",
                                                "rocfft",    "onemkl",        "hip_native"};
        const auto i = static_cast<std::size_t>(value);
        return i < sizeof(names) / sizeof(names[0]) ? names[i] : "unknown";
    }

    std::vector<std::string> provider_search_paths()
    {
        std::vector<std::string> dirs;
        if (const char *env = std::getenv("SYCLFFT_PLUGIN_PATH")) {
            std::string_view rest(env);
            while (!rest.empty()) {
                const auto colon = rest.find(':');
                const auto entry = rest.substr(0, colon);
                if (!entry.empty()) {
                    dirs.emplace_back(entry);
                }
                rest = colon == std::string_view::npos ? std::string_view{} : rest.substr(colon + 1);
            }
        }
        const std::string here = own_directory();
        if (!here.empty()) {
            dirs.push_back(here + "/syclfft/providers");
        }
        return dirs;
    }

    loaded_provider_symbol load_provider_symbol(const std::string &provider_file, const std::string &symbol_name)
    {
        std::string attempts;
        for (const std::string &dir : provider_search_paths()) {
            const std::string path = dir + "/" + provider_file + ".so";
            std::string why;
            std::shared_ptr<void> module = open_module(path, why);
            if (module) {
                if (void *entry = dlsym(module.get(), symbol_name.c_str())) {
                    trace(("core: " + symbol_name + " found in " + path).c_str());
                    return {entry, std::move(module)};
                }
                why = "no entry point " + symbol_name;
            }
            attempts += (attempts.empty() ? "" : "; ") + path + " (" + why + ")";
        }
        if (attempts.empty()) {
            attempts = "no search directory";
        }
        throw exception(error_code::provider_unavailable,
                        "provider module " + provider_file + " is not usable: " + attempts);
    }

    loaded_host_provider load_host_provider(const std::string &provider_file)
    {
        loaded_provider_symbol found = load_provider_symbol(provider_file, "syclfft_get_host_provider_v1");
        const auto get_table = reinterpret_cast<syclfft_get_host_provider_v1_fn>(found.symbol);
        const syclfft_host_provider_v1 *table = get_table();
        if (table == nullptr) {
            throw exception(error_code::plugin_error, "provider module " + provider_file + " returned no host table");
        }
        if (table->abi_version != SYCLFFT_PROVIDER_ABI_VERSION) {
            throw exception(error_code::plugin_error,
                            "provider module " + provider_file + " has host ABI " + std::to_string(table->abi_version)
                                + ", expected " + std::to_string(SYCLFFT_PROVIDER_ABI_VERSION));
        }
        return {table, std::move(found.module)};
    }

} // namespace syclfft::detail
