// host::plan: synchronous transforms of host memory, run by a host-provider module behind
// the ABI v1 function table (module `syclfft_provider_fftw`).
//
// Behaviour contract of the reference's host front-end (src/host/host.cpp:37-215), which
// this file implements with its own structure:
//   construction  shape checks (rank 1..3, non-zero lengths and batch, no overflow) and the
//                 direction -> invalid_argument; a preferred provider other than automatic
//                 or fftw -> provider_unavailable; module missing -> provider_unavailable,
//                 wrong ABI -> plugin_error; the module's create() failing -> planning_failed
//   execute       moved-from plan -> invalid_state; a null buffer -> invalid_pointer; the
//                 call form must match the placement -> invalid_argument; executions of
//                 one plan are serialized; a failing provider execute() -> execution_failed;
//                 the normalization factor is applied on the host after the transform
//                 (the ABI carries no normalization).
#include <algorithm>
#include <array>
#include <mutex>
#include <syclfft/detail/provider_loader.hpp>
#include <syclfft/detail/trace.hpp>
#include <syclfft/host.hpp>

namespace syclfft::host
{
    namespace
    {
        constexpr const char *k_module_file = "syclfft_provider_fftw";

        // Error text buffer handed across the module boundary.
        struct module_error
        {
            std::array<char, 1024> text{};

            std::string or_else(const char *fallback) const
            {
                return text[0] != '\0' ? std::string(text.data()) : std::string(fallback);
            }
        };

        // One plan inside a provider module: the table it came from, its opaque handle, and
        // a share of the module so the library outlives every handle it created.
        class module_plan
        {
        public:
            module_plan(detail::loaded_host_provider loaded, const syclfft_host_plan_config_v1 &cfg)
              : table_(loaded.api)
              , library_(std::move(loaded.module))
            {
                module_error e;
                handle_ = table_->create(&cfg, e.text.data(), e.text.size());
                if (handle_ == nullptr) {
                    throw exception(error_code::planning_failed, e.or_else("host provider could not create the plan"));
                }
            }
            module_plan(const module_plan &) = delete;
            module_plan &operator=(const module_plan &) = delete;
            ~module_plan()
            {
                table_->destroy(handle_);
            }

            void run(const void *in, void *out) const
            {
                module_error e;
                // the ABI's execute takes a mutable input pointer; providers only read it
                if (table_->execute(handle_, const_cast<void *>(in), out, e.text.data(), e.text.size()) != 0) {
                    throw exception(error_code::execution_failed, e.or_else("host provider execution failed"));
                }
            }

        private:
            const syclfft_host_provider_v1 *table_;
            std::shared_ptr<void> library_;
            void *handle_{};
        };

        template <class Scalar>
        syclfft_host_plan_config_v1 module_config(
            const std::vector<std::size_t> &lengths, std::size_t batch, syclfft::direction dir,
            const plan_options &opts)
        {
            syclfft_host_plan_config_v1 cfg{};
            cfg.abi_version = SYCLFFT_PROVIDER_ABI_VERSION;
            cfg.scalar = sizeof(Scalar) == sizeof(double) ? SYCLFFT_SCALAR_DOUBLE_V1 : SYCLFFT_SCALAR_FLOAT_V1;
            cfg.direction = static_cast<int32_t>(dir); // forward = -1, backward = +1 in both
            cfg.rank = static_cast<uint32_t>(lengths.size());
            std::copy(lengths.begin(), lengths.end(), cfg.lengths);
            cfg.batch_count = batch;
            cfg.in_place = opts.placement == placement::in_place;
            cfg.measure = opts.planning == planning_mode::measure;
            return cfg;
        }

        bool host_provider_allowed(provider p)
        {
            return p == provider::automatic || p == provider::fftw;
        }

        // How the caller invoked execute: execute(inout) or execute(in, out).
        enum class call_form
        {
            inout,
            in_out
        };

        void check_call(placement where, const void *in, const void *out, call_form form)
        {
            if (in == nullptr || out == nullptr) {
                throw exception(error_code::invalid_pointer, "host plan: input and output must not be null");
            }
            if (where == placement::in_place && form != call_form::inout) {
                throw exception(error_code::invalid_argument, "host plan: an in-place plan is executed with execute(inout)");
            }
            if (where == placement::out_of_place && form != call_form::in_out) {
                throw exception(
                    error_code::invalid_argument, "host plan: an out-of-place plan needs separate input and output");
            }
        }
    } // namespace

    template <class Scalar>
    class plan<Scalar>::impl
    {
    public:
        impl(std::vector<std::size_t> lengths, std::size_t batch, syclfft::direction dir, plan_options opts)
          : shape(std::move(lengths))
          , batch_count(batch)
          , points(detail::checked_element_count(shape, batch_count))
          , dir(dir)
          , options(opts)
          , factor(static_cast<Scalar>(detail::normalization_scale(opts.normalization, dir, detail::transform_size(shape))))
          , backend(open_backend())
        {}

        void execute(const complex_type *in, complex_type *out, call_form form)
        {
            check_call(options.placement, in, out, form);
            std::lock_guard<std::mutex> serial(order);
            backend.run(in, out);
            if (factor != Scalar{1}) {
                std::for_each(out, out + points, [f = factor](complex_type &v) { v *= f; });
            }
        }

        std::vector<std::size_t> shape;
        std::size_t batch_count;
        std::size_t points;
        syclfft::direction dir;
        plan_options options;
        Scalar factor;

    private:
        module_plan open_backend() const
        {
            if (dir != syclfft::direction::forward && dir != syclfft::direction::backward) {
                throw exception(error_code::invalid_argument, "host plan: direction must be forward or backward");
            }
            if (!host_provider_allowed(options.preferred_provider)) {
                throw exception(error_code::provider_unavailable,
                                std::string("host plan: provider ") + detail::provider_name(options.preferred_provider)
                                    + " does not run host plans (automatic or fftw only)");
            }
            detail::trace("host: opening the host provider module");
            return module_plan(detail::load_host_provider(k_module_file),
                               module_config<Scalar>(shape, batch_count, dir, options));
        }

        module_plan backend;
        std::mutex order;
    };

    namespace
    {
        template <class Impl>
        Impl &live(const std::unique_ptr<Impl> &p)
        {
            if (!p) {
                throw exception(error_code::invalid_state, "Cannot execute a moved-from plan");
            }
            return *p;
        }
    } // namespace

    template <class Scalar>
    plan<Scalar>::plan(std::vector<std::size_t> lengths, std::size_t batch_count, syclfft::direction dir,
                       plan_options options)
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
        live(impl_).execute(inout, inout, call_form::inout);
    }
    template <class Scalar>
    void plan<Scalar>::execute(const complex_type *input, complex_type *output)
    {
        live(impl_).execute(input, output, call_form::in_out);
    }
    template <class Scalar>
    syclfft::span<const std::size_t> plan<Scalar>::shape() const noexcept
    {
        if (!impl_) {
            return {};
        }
        return syclfft::span<const std::size_t>{impl_->shape};
    }
    template <class Scalar>
    std::size_t plan<Scalar>::batch_count() const noexcept
    {
        return impl_ ? impl_->batch_count : 0;
    }
    template <class Scalar>
    syclfft::direction plan<Scalar>::direction() const noexcept
    {
        return impl_ ? impl_->dir : syclfft::direction::forward;
    }
    template <class Scalar>
    syclfft::provider plan<Scalar>::selected_provider() const noexcept
    {
        return impl_ ? provider::fftw : provider::automatic;
    }
    template <class Scalar>
    std::size_t plan<Scalar>::scratch_size_bytes() const noexcept
    {
        // the provider's working set: the output array, plus the input for out-of-place plans
        if (!impl_) {
            return 0;
        }
        const std::size_t arrays = impl_->options.placement == placement::in_place ? 1 : 2;
        return arrays * impl_->points * sizeof(complex_type);
    }

    template class plan<float>;
    template class plan<double>;

} // namespace syclfft::host
