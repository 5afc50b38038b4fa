// Host provider module `syclfft_provider_fftw` (host ABI v1, include/syclfft/detail/provider_abi.h)
// over an FFTW3-API library loaded at run time: libfftw3 / libfftw3f when installed, else
// Intel MKL's FFTW3 interface (libmkl_rt).  Plan layout follows the reference's FFTW
// provider (src/providers/fftw/fftw_provider.cpp:45-63, 141-244): plan_many_dft over
// rank/lengths, howmany = batch, unit stride, distance = prod(lengths), sign from the
// direction, ESTIMATE or MEASURE; plans are created under one global planner mutex; no
// normalization here (host::plan applies it).  Plans are made FFTW_UNALIGNED so any
// caller pointers are valid with the new-array execute functions.
#include <dlfcn.h>

#include <array>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <syclfft/detail/provider_abi.h>
#include <syclfft/detail/provider_error_buffer.hpp>

namespace
{
    constexpr unsigned k_measure = 0u;
    constexpr unsigned k_estimate = 1u << 6;
    constexpr unsigned k_unaligned = 1u << 1;

    using plan_many_fn = void *(*)(int, const int *, int, void *, const int *, int, int, void *, const int *, int,
                                   int, int, unsigned);
    using exec_fn = void (*)(void *, void *, void *);
    using destroy_fn = void (*)(void *);
    using malloc_fn = void *(*)(size_t);
    using free_fn = void (*)(void *);

    struct fftw_api
    {
        plan_many_fn plan_many{};
        exec_fn execute{};
        destroy_fn destroy{};
        malloc_fn alloc{};
        free_fn release{};
    };

    struct library
    {
        void *handle{};
        std::string name;
        fftw_api f32, f64;
        std::string error;
    };

    std::mutex &planner_mutex()
    {
        static std::mutex m;
        return m;
    }

    template <class T>
    bool resolve(void *h, const char *prefix, T &api)
    {
        auto sym = [&](const char *suffix) { return dlsym(h, (std::string(prefix) + suffix).c_str()); };
        api.plan_many = reinterpret_cast<plan_many_fn>(sym("_plan_many_dft"));
        api.execute = reinterpret_cast<exec_fn>(sym("_execute_dft"));
        api.destroy = reinterpret_cast<destroy_fn>(sym("_destroy_plan"));
        api.alloc = reinterpret_cast<malloc_fn>(sym("_malloc"));
        api.release = reinterpret_cast<free_fn>(sym("_free"));
        return api.plan_many && api.execute && api.destroy && api.alloc && api.release;
    }

    const library &lib()
    {
        static library L = [] {
            library out;
            if (const char *forced = std::getenv("SYCLFFT_FFTW_LIBRARY")) {
                out.handle = dlopen(forced, RTLD_NOW | RTLD_LOCAL);
                out.name = forced;
            }
            if (!out.handle) {
                void *d = dlopen("libfftw3.so.3", RTLD_NOW | RTLD_LOCAL);
                void *f = dlopen("libfftw3f.so.3", RTLD_NOW | RTLD_LOCAL);
                if (d && f && resolve(d, "fftw", out.f64) && resolve(f, "fftwf", out.f32)) {
                    out.handle = d;
                    out.name = "fftw3";
                    return out;
                }
            }
            for (const char *cand : {"libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"}) {
                if (out.handle) {
                    break;
                }
                out.handle = dlopen(cand, RTLD_NOW | RTLD_LOCAL);
                out.name = "mkl-fftw3 (" + std::string(cand) + ")";
            }
            if (!out.handle) {
                out.error = "no FFTW3-API library found (libfftw3 or libmkl_rt)";
                return out;
            }
            if (!resolve(out.handle, "fftw", out.f64) || !resolve(out.handle, "fftwf", out.f32)) {
                out.error = "library lacks the FFTW3 plan_many_dft API";
                out.handle = nullptr;
            }
            return out;
        }();
        return L;
    }

    struct holder
    {
        const fftw_api *api{};
        void *plan{};
    };

    void *create(const syclfft_host_plan_config_v1 *cfg, char *err, size_t cap)
    {
        try {
            if (!cfg || cfg->abi_version != SYCLFFT_PROVIDER_ABI_VERSION) {
                throw std::runtime_error("incompatible host plan config");
            }
            const library &L = lib();
            if (!L.handle) {
                throw std::runtime_error(L.error);
            }
            if (cfg->rank < 1 || cfg->rank > 3 || cfg->batch_count == 0) {
                throw std::runtime_error("invalid rank or batch");
            }
            std::array<int, 3> n{};
            uint64_t dist = 1;
            for (uint32_t i = 0; i < cfg->rank; ++i) {
                if (cfg->lengths[i] == 0 || cfg->lengths[i] > 0x7fffffffull) {
                    throw std::runtime_error("FFT length does not fit FFTW's int");
                }
                n[i] = static_cast<int>(cfg->lengths[i]);
                dist *= cfg->lengths[i];
            }
            if (dist > 0x7fffffffull || cfg->batch_count > 0x7fffffffull) {
                throw std::runtime_error("FFT size does not fit FFTW's int");
            }
            const bool dbl = cfg->scalar == SYCLFFT_SCALAR_DOUBLE_V1;
            if (!dbl && cfg->scalar != SYCLFFT_SCALAR_FLOAT_V1) {
                throw std::runtime_error("unsupported scalar");
            }
            const fftw_api &api = dbl ? L.f64 : L.f32;
            const size_t bytes = dist * cfg->batch_count * (dbl ? 16 : 8);
            void *a = api.alloc(bytes);
            void *b = cfg->in_place ? a : api.alloc(bytes);
            if (!a || !b) {
                throw std::bad_alloc();
            }
            const unsigned flags = (cfg->measure ? k_measure : k_estimate) | k_unaligned;
            void *p;
            {
                std::scoped_lock lock(planner_mutex());
                p = api.plan_many(static_cast<int>(cfg->rank), n.data(), static_cast<int>(cfg->batch_count), a,
                                  nullptr, 1, static_cast<int>(dist), b, nullptr, 1, static_cast<int>(dist),
                                  cfg->direction, flags);
            }
            api.release(a);
            if (b != a) {
                api.release(b);
            }
            if (!p) {
                throw std::runtime_error("FFTW planner returned no plan");
            }
            syclfft::detail::set_error_buffer(err, cap, "");
            return new holder{&api, p};
        } catch (const std::exception &ex) {
            syclfft::detail::set_error_buffer(err, cap, ex.what());
        } catch (...) {
            syclfft::detail::set_error_buffer(err, cap, "unknown FFTW provider error");
        }
        return nullptr;
    }

    void destroy(void *h)
    {
        auto *p = static_cast<holder *>(h);
        if (p) {
            std::scoped_lock lock(planner_mutex());
            p->api->destroy(p->plan);
            delete p;
        }
    }

    int execute(void *h, void *in, void *out, char *err, size_t cap)
    {
        auto *p = static_cast<holder *>(h);
        if (!p || !in || !out) {
            syclfft::detail::set_error_buffer(err, cap, "null plan or buffer");
            return 1;
        }
        p->api->execute(p->plan, in, out);
        return 0;
    }

    const syclfft_host_provider_v1 k_provider = {SYCLFFT_PROVIDER_ABI_VERSION, "fftw", create, destroy, execute};
} // namespace

SYCLFFT_PROVIDER_EXPORT const syclfft_host_provider_v1 *syclfft_get_host_provider_v1(void)
{
    return &k_provider;
}

// Which FFTW3-API library backs this module ("fftw3" or "mkl-fftw3 (...)"); empty if none.
extern "C" __attribute__((visibility("default"))) const char *syclfft_fftw_provider_backend(void)
{
    const auto &L = lib();
    return L.handle ? L.name.c_str() : "";
}
