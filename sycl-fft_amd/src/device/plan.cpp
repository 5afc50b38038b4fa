// Device plan front-end over the MI355X provider module (C ABI in
// include/syclfft_hip_provider.h), replacing the reference's SYCL plan implementation
// (src/sycl/sycl.cpp:405-947).  Kept from the reference: validation and error codes,
// provider selection and diagnostics, executions of one plan serialized through the
// previous execution's event, dependency events, moved-from plans, query_providers.
#include <hip/hip_runtime_api.h>

#include <array>
#include <mutex>
#include <optional>
#include <string>
#include <syclfft/detail/provider_loader.hpp>
#include <syclfft/detail/trace.hpp>
#include <syclfft/syclfft.hpp>

#include "syclfft_hip_provider.h"

namespace syclfft
{
    namespace
    {
        void hip_or_throw(hipError_t st, error_code code, const char *what)
        {
            if (st != hipSuccess) {
                throw exception(code, std::string(what) + ": " + hipGetErrorString(st));
            }
        }

        error_code code_of(int status)
        {
            switch (status) {
            case SYCLFFT_HIP_STATUS_INVALID_ARGUMENT:
                return error_code::invalid_argument;
            case SYCLFFT_HIP_STATUS_INVALID_STATE:
                return error_code::invalid_state;
            case SYCLFFT_HIP_STATUS_INVALID_POINTER:
                return error_code::invalid_pointer;
            case SYCLFFT_HIP_STATUS_UNSUPPORTED_PRECISION:
                return error_code::unsupported_precision;
            case SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE:
                return error_code::provider_unavailable;
            case SYCLFFT_HIP_STATUS_PLANNING_FAILED:
                return error_code::planning_failed;
            case SYCLFFT_HIP_STATUS_PLUGIN_ERROR:
                return error_code::plugin_error;
            default:
                return error_code::execution_failed;
            }
        }

        struct loaded_device_provider
        {
            const syclfft_device_provider_v1 *api{};
            std::shared_ptr<void> module;
        };

        // Load syclfft_provider_hip and check its ABI (the role of try_load_cufft,
        // src/sycl/sycl.cpp:347-401).
        loaded_device_provider try_load_hip(int device, bool required, std::string *reason)
        {
            try {
                auto sym = detail::load_provider_symbol("syclfft_provider_hip", "syclfft_get_device_provider_v1");
                const auto getter = reinterpret_cast<syclfft_get_device_provider_v1_fn>(sym.symbol);
                const auto *api = getter();
                if (!api || api->abi_version != SYCLFFT_DEVICE_PROVIDER_ABI_VERSION) {
                    throw exception(error_code::plugin_error, "HIP provider ABI mismatch");
                }
                std::array<char, 1024> err{};
                if (!api->supports(device, err.data(), err.size())) {
                    if (reason) {
                        *reason = err.data();
                    }
                    if (required) {
                        throw exception(error_code::provider_unavailable, err[0] ? err.data() : "HIP provider rejected the device");
                    }
                    return {};
                }
                return {api, std::move(sym.module)};
            } catch (const exception &ex) {
                if (reason) {
                    *reason = ex.what();
                }
                if (required) {
                    throw;
                }
                return {};
            }
        }
    } // namespace

    // ------------------------------------------------------------------ queue / event

    queue::queue()
    {
        if (hipGetDevice(&device_) != hipSuccess) {
            device_ = 0;
        }
    }

    queue::queue(int device, void *stream)
      : device_(device)
      , stream_(stream)
    {}

    void queue::wait() const
    {
        int saved = 0;
        (void)hipGetDevice(&saved);
        hip_or_throw(hipSetDevice(device_), error_code::execution_failed, "hipSetDevice");
        const hipError_t st = hipStreamSynchronize(static_cast<hipStream_t>(stream_));
        (void)hipSetDevice(saved);
        hip_or_throw(st, error_code::execution_failed, "hipStreamSynchronize");
    }

    class event_access
    {
    public:
        static event make(hipEvent_t e)
        {
            event out;
            out.handle_ = std::shared_ptr<void>(e, [](void *h) {
                if (h) {
                    (void)hipEventDestroy(static_cast<hipEvent_t>(h));
                }
            });
            return out;
        }
    };

    event event::record(const queue &q)
    {
        int saved = 0;
        (void)hipGetDevice(&saved);
        hip_or_throw(hipSetDevice(q.device()), error_code::execution_failed, "hipSetDevice");
        hipEvent_t e{};
        hip_or_throw(hipEventCreateWithFlags(&e, hipEventDisableTiming), error_code::execution_failed, "hipEventCreate");
        const hipError_t st = hipEventRecord(e, static_cast<hipStream_t>(q.native_stream()));
        (void)hipSetDevice(saved);
        if (st != hipSuccess) {
            (void)hipEventDestroy(e);
            hip_or_throw(st, error_code::execution_failed, "hipEventRecord");
        }
        return event_access::make(e);
    }

    void event::wait() const
    {
        if (handle_) {
            hip_or_throw(hipEventSynchronize(static_cast<hipEvent_t>(handle_.get())), error_code::execution_failed,
                         "hipEventSynchronize");
        }
    }

    void *event::native_handle() const noexcept
    {
        return handle_.get();
    }

    bool event::complete() const
    {
        return !handle_ || hipEventQuery(static_cast<hipEvent_t>(handle_.get())) == hipSuccess;
    }

    // ------------------------------------------------------------------ plan

    template <class Scalar>
    class plan<Scalar>::impl
    {
    public:
        impl(queue q, std::vector<std::size_t> lengths, std::size_t batch_count, syclfft::direction dir, plan_options opts)
          : queue_(q)
          , lengths_(std::move(lengths))
          , batch_count_(batch_count)
          , element_count_(detail::checked_element_count(lengths_, batch_count_))
          , direction_(dir)
          , options_(opts)
        {
            if (dir != syclfft::direction::forward && dir != syclfft::direction::backward) {
                throw exception(error_code::invalid_argument, "Invalid FFT direction");
            }
            switch (options_.preferred_provider) {
            case provider::automatic:
            case provider::portable_sycl: // the portable path's kernels are the HIP provider here
            case provider::hip_native:
                break;
            case provider::fftw:
                throw exception(error_code::provider_unavailable, "FFTW can only be selected for a host plan (syclfft::host)");
            case provider::cufft:
                throw exception(error_code::provider_unavailable, "cuFFT requires a CUDA device");
            default:
                throw exception(
                    error_code::provider_unavailable,
                    std::string(detail::provider_name(options_.preferred_provider)) + " provider is not implemented");
            }
            provider_ = try_load_hip(queue_.device(), true, nullptr);
            syclfft_hip_plan_config_v1 cfg{};
            cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
            cfg.scalar = std::is_same_v<Scalar, float> ? SYCLFFT_HIP_SCALAR_FLOAT : SYCLFFT_HIP_SCALAR_DOUBLE;
            cfg.direction = static_cast<int32_t>(dir);
            cfg.rank = static_cast<uint32_t>(lengths_.size());
            for (std::size_t i = 0; i < lengths_.size(); ++i) {
                cfg.lengths[i] = lengths_[i];
            }
            cfg.batch_count = batch_count_;
            cfg.in_place = options_.placement == placement::in_place ? 1u : 0u;
            cfg.normalization = static_cast<uint32_t>(options_.normalization);
            cfg.measure = options_.planning == planning_mode::measure ? 1u : 0u;
            cfg.device = queue_.device();
            std::array<char, 1024> err{};
            handle_ = provider_.api->create(&cfg, err.data(), err.size());
            if (!handle_) {
                throw exception(error_code::planning_failed, err[0] ? err.data() : "HIP plan creation failed");
            }
        }

        impl(const impl &) = delete;
        impl &operator=(const impl &) = delete;

        ~impl()
        {
            try {
                if (last_) {
                    last_->wait();
                }
            } catch (...) {
            }
            if (handle_) {
                provider_.api->destroy(handle_);
            }
        }

        void validate_pointer(const void *p, const char *name) const
        {
            if (!p) {
                throw exception(error_code::invalid_pointer, std::string(name) + " must be non-null");
            }
            hipPointerAttribute_t attr{};
            if (hipPointerGetAttributes(&attr, p) != hipSuccess
                || (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)) {
                (void)hipGetLastError();
                throw exception(error_code::invalid_pointer, std::string(name) + " is not a device allocation of the plan's device");
            }
        }

        event execute(const complex_type *in, complex_type *out, syclfft::span<const event> deps, bool one_pointer)
        {
            validate_pointer(in, "input");
            validate_pointer(out, "output");
            if (options_.placement == placement::in_place && !one_pointer) {
                throw exception(error_code::invalid_argument, "In-place plan must be executed with execute(inout)");
            }
            if (options_.placement == placement::out_of_place && one_pointer) {
                throw exception(error_code::invalid_argument, "Out-of-place plan requires input and output arguments");
            }
            if (options_.placement == placement::out_of_place && static_cast<const void *>(in) == out) {
                throw exception(error_code::invalid_argument, "Out-of-place plan requires distinct allocations");
            }
            std::scoped_lock lock(mutex_);
            std::vector<void *> handles;
            for (const auto &d : deps) {
                if (d.native_handle()) {
                    handles.push_back(d.native_handle());
                }
            }
            if (last_ && last_->native_handle()) {
                handles.push_back(last_->native_handle());
            }
            int saved = 0;
            (void)hipGetDevice(&saved);
            hip_or_throw(hipSetDevice(queue_.device()), error_code::execution_failed, "hipSetDevice");
            hipEvent_t done{};
            hip_or_throw(hipEventCreateWithFlags(&done, hipEventDisableTiming), error_code::execution_failed,
                         "hipEventCreate");
            event result = event_access::make(done);
            std::array<char, 1024> err{};
            const int st = provider_.api->execute(handle_, in, out, queue_.native_stream(), handles.data(),
                                                  handles.size(), done, err.data(), err.size());
            (void)hipSetDevice(saved);
            if (st != SYCLFFT_HIP_STATUS_OK) {
                throw exception(code_of(st), err[0] ? err.data() : "HIP execution failed");
            }
            last_ = result;
            return result;
        }

        queue queue_;
        std::vector<std::size_t> lengths_;
        std::size_t batch_count_{};
        std::size_t element_count_{};
        syclfft::direction direction_{};
        plan_options options_{};
        loaded_device_provider provider_;
        void *handle_{};
        std::optional<event> last_;
        std::mutex mutex_;
    };

    template <class Scalar>
    plan<Scalar>::plan(queue q, std::vector<std::size_t> lengths, std::size_t batch_count, syclfft::direction dir,
                       plan_options options)
      : impl_(std::make_unique<impl>(q, std::move(lengths), batch_count, dir, options))
    {}
    template <class Scalar>
    plan<Scalar>::plan(plan &&) noexcept = default;
    template <class Scalar>
    plan<Scalar> &plan<Scalar>::operator=(plan &&) noexcept = default;
    template <class Scalar>
    plan<Scalar>::~plan() = default;

#define SYCLFFT_REQUIRE_IMPL                                                                      \
    if (!impl_) {                                                                                 \
        throw exception(error_code::invalid_state, "Cannot execute a moved-from plan");          \
    }

    template <class Scalar>
    event plan<Scalar>::execute(complex_type *inout)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(inout, inout, {}, true);
    }
    template <class Scalar>
    event plan<Scalar>::execute(complex_type *inout, const event &dependency)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(inout, inout, syclfft::span<const event>{&dependency, 1}, true);
    }
    template <class Scalar>
    event plan<Scalar>::execute(complex_type *inout, syclfft::span<const event> dependencies)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(inout, inout, dependencies, true);
    }
    template <class Scalar>
    event plan<Scalar>::execute(const complex_type *in, complex_type *out)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(in, out, {}, false);
    }
    template <class Scalar>
    event plan<Scalar>::execute(const complex_type *in, complex_type *out, const event &dependency)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(in, out, syclfft::span<const event>{&dependency, 1}, false);
    }
    template <class Scalar>
    event plan<Scalar>::execute(const complex_type *in, complex_type *out, syclfft::span<const event> dependencies)
    {
        SYCLFFT_REQUIRE_IMPL
        return impl_->execute(in, out, dependencies, false);
    }
#undef SYCLFFT_REQUIRE_IMPL

    template <class Scalar>
    syclfft::span<const std::size_t> plan<Scalar>::shape() const noexcept
    {
        return impl_ ? syclfft::span<const std::size_t>{impl_->lengths_} : syclfft::span<const std::size_t>{};
    }
    template <class Scalar>
    std::size_t plan<Scalar>::batch_count() const noexcept
    {
        return impl_ ? impl_->batch_count_ : 0;
    }
    template <class Scalar>
    syclfft::direction plan<Scalar>::direction() const noexcept
    {
        return impl_ ? impl_->direction_ : syclfft::direction::forward;
    }
    template <class Scalar>
    syclfft::provider plan<Scalar>::selected_provider() const noexcept
    {
        return impl_ ? provider::hip_native : provider::automatic;
    }
    template <class Scalar>
    std::size_t plan<Scalar>::scratch_size_bytes() const noexcept
    {
        return impl_ ? static_cast<std::size_t>(impl_->provider_.api->scratch_bytes(impl_->handle_)) : 0;
    }

    std::vector<provider_status> query_providers(const queue &q)
    {
        std::vector<provider_status> out;
        std::string reason;
        const auto hip = try_load_hip(q.device(), false, &reason);
        out.push_back({provider::hip_native, "hip_native", true, hip.api != nullptr, hip.api ? "" : reason});
        out.push_back({provider::portable_sycl, "portable_sycl", true, hip.api != nullptr,
                       hip.api ? "served by the hip_native kernels" : reason});
        out.push_back({provider::cufft, "cufft", false, false, "provider was not built (NVIDIA only)"});
        std::string fftw_reason = "queue is not a CPU device (use syclfft::host plans)";
        bool fftw_built = true;
        try {
            (void)detail::load_host_provider("syclfft_provider_fftw");
        } catch (const std::exception &ex) {
            fftw_built = false;
            fftw_reason = ex.what();
        }
        out.push_back({provider::fftw, "fftw", fftw_built, false, fftw_reason});
        out.push_back({provider::rocfft, "rocfft", false, false, "reserved for a future provider"});
        out.push_back({provider::onemkl, "onemkl", false, false, "reserved for a future provider"});
        return out;
    }

    template class plan<float>;
    template class plan<double>;

} // namespace syclfft
