// host::plan: synchronous plans on host memory through a host-provider module (ABI v1,
// module `syclfft_provider_fftw`).  Semantics of the reference's src/host/host.cpp:37-215:
// provider must be automatic or fftw, config built from the plan, execute checks pointers
// and placement, serializes on a mutex, and applies the normalization on the host.
#include <array>
#include <mutex>
#include <syclfft/detail/provider_loader.hpp>
#include <syclfft/detail/trace.hpp>
#include <syclfft/host.hpp>

namespace syclfft::host
{
    template <class Scalar>
    class plan<Scalar>::impl
    {
    public:
        impl(std::vector<std::size_t> lengths, std::size_t batch_count, syclfft::direction dir, plan_options opts)
          : lengths_(std::move(lengths))
          , batch_count_(batch_count)
          , element_count_(detail::checked_element_count(lengths_, batch_count_))
          , direction_(dir)
          , options_(opts)
        {
            if (dir != syclfft::direction::forward && dir != syclfft::direction::backward) {
                throw exception(error_code::invalid_argument, "Invalid FFT direction");
            }
            if (options_.preferred_provider != provider::automatic && options_.preferred_provider != provider::fftw) {
                throw exception(error_code::provider_unavailable, "Host plans only support the FFTW provider");
            }
            detail::trace("host: loading FFTW provider");
            auto loaded = detail::load_host_provider("syclfft_provider_fftw");
            api_ = loaded.api;
            module_ = std::move(loaded.module);

            syclfft_host_plan_config_v1 cfg{};
            cfg.abi_version = SYCLFFT_PROVIDER_ABI_VERSION;
            cfg.scalar = std::is_same_v<Scalar, float> ? SYCLFFT_SCALAR_FLOAT_V1 : SYCLFFT_SCALAR_DOUBLE_V1;
            cfg.direction = dir == syclfft::direction::forward ? SYCLFFT_DIRECTION_FORWARD_V1 : SYCLFFT_DIRECTION_BACKWARD_V1;
            cfg.rank = static_cast<uint32_t>(lengths_.size());
            cfg.batch_count = batch_count_;
            cfg.in_place = options_.placement == placement::in_place ? 1u : 0u;
            cfg.measure = options_.planning == planning_mode::measure ? 1u : 0u;
            for (std::size_t i = 0; i < lengths_.size(); ++i) {
                cfg.lengths[i] = lengths_[i];
            }
            std::array<char, 1024> err{};
            handle_ = api_->create(&cfg, err.data(), err.size());
            if (!handle_) {
                throw exception(error_code::planning_failed, err[0] ? err.data() : "FFTW plan creation failed");
            }
        }
        impl(const impl &) = delete;
        impl &operator=(const impl &) = delete;
        ~impl()
        {
            if (handle_) {
                api_->destroy(handle_);
            }
        }

        void execute(const complex_type *in, complex_type *out, bool one_pointer)
        {
            if (!in || !out) {
                throw exception(error_code::invalid_pointer, "FFT input and output must be non-null");
            }
            if (options_.placement == placement::in_place && !one_pointer) {
                throw exception(error_code::invalid_argument, "In-place plan must be executed with execute(inout)");
            }
            if (options_.placement == placement::out_of_place && one_pointer) {
                throw exception(
                    error_code::invalid_argument, "Out-of-place plan requires distinct input and output arguments");
            }
            std::scoped_lock lock(mutex_);
            std::array<char, 1024> err{};
            if (api_->execute(handle_, const_cast<complex_type *>(in), out, err.data(), err.size()) != 0) {
                throw exception(error_code::execution_failed, err[0] ? err.data() : "FFTW execution failed");
            }
            const auto scale = static_cast<Scalar>(
                detail::normalization_scale(options_.normalization, direction_, detail::transform_size(lengths_)));
            if (scale != Scalar{1}) {
                for (std::size_t i = 0; i < element_count_; ++i) {
                    out[i] *= scale;
                }
            }
        }

        std::vector<std::size_t> lengths_;
        std::size_t batch_count_{};
        std::size_t element_count_{};
        syclfft::direction direction_{};
        plan_options options_{};
        const syclfft_host_provider_v1 *api_{};
        std::shared_ptr<void> module_;
        void *handle_{};
        std::mutex mutex_;
    };

    template <class Scalar>
    plan<Scalar>::plan(std::vector<std::size_t> lengths, std::size_t batch_count, syclfft::direction dir, plan_options options)
      : impl_(std::make_unique<impl>(std::move(lengths), batch_count, dir, options))
    {}
    template <class Scalar>
    plan<Scalar>::plan(plan &&) noexcept = default;
    template <class Scalar>
    plan<Scalar> &plan<Scalar>::operator=(plan &&) noexcept = default;
    template <class Scalar>
    plan<Scalar>::~plan() = default;

    template <class Scalar>
    void plan<Scalar>::execute(complex_type *inout)
    {
        if (!impl_) {
            throw exception(error_code::invalid_state, "Cannot execute a moved-from plan");
        }
        impl_->execute(inout, inout, true);
    }
    template <class Scalar>
    void plan<Scalar>::execute(const complex_type *in, complex_type *out)
    {
        if (!impl_) {
            throw exception(error_code::invalid_state, "Cannot execute a moved-from plan");
        }
        impl_->execute(in, out, false);
    }
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
        return impl_ ? provider::fftw : provider::automatic;
    }
    template <class Scalar>
    std::size_t plan<Scalar>::scratch_size_bytes() const noexcept
    {
        if (!impl_) {
            return 0;
        }
        const std::size_t buffers = impl_->options_.placement == placement::in_place ? 1 : 2;
        return buffers * impl_->element_count_ * sizeof(complex_type);
    }

    template class plan<float>;
    template class plan<double>;

} // namespace syclfft::host
