// Device plan API on HIP (MI355X), the drop-in for the reference's SYCL plan API
// (reference: include/syclfft/syclfft.hpp:39-129).  sycl::queue and sycl::event are
// replaced by syclfft::queue (a HIP device + stream) and syclfft::event (a shared HIP
// event); plan construction, the six execute overloads, the accessors and the plan_*
// helpers keep the reference's signatures and semantics (serialized executions of one
// plan, placement checks, dependency events, moved-from plans throw invalid_state).
#pragma once

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <syclfft/common.hpp>
#include <type_traits>
#include <vector>

namespace syclfft
{

    // Interleaved {real, imag} POD, layout-compatible with std::complex<Scalar>
    // (the layout the reference asserts at tests/sycl_tests.cpp:268-287).
    template <class Scalar>
    struct complex
    {
        Scalar re{};
        Scalar im{};

        constexpr complex() = default;
        constexpr complex(Scalar r, Scalar i = Scalar{})
          : re(r)
          , im(i)
        {}
        constexpr complex(const std::complex<Scalar> &v)
          : re(v.real())
          , im(v.imag())
        {}
        constexpr operator std::complex<Scalar>() const
        {
            return {re, im};
        }
        [[nodiscard]] constexpr Scalar real() const
        {
            return re;
        }
        [[nodiscard]] constexpr Scalar imag() const
        {
            return im;
        }
    };

    // A HIP device and one of its streams (the role of sycl::queue).  The default queue is
    // the current device's null stream.
    class SYCLFFT_EXPORT queue
    {
    public:
        queue();                                      // current device, null stream
        explicit queue(int device, void *stream = nullptr); // stream: a hipStream_t
        [[nodiscard]] int device() const noexcept
        {
            return device_;
        }
        [[nodiscard]] void *native_stream() const noexcept
        {
            return stream_;
        }
        void wait() const; // hipStreamSynchronize

    private:
        int device_{};
        void *stream_{};
    };

    // Completion handle of an enqueued operation (the role of sycl::event): a shared HIP
    // event recorded on the queue's stream.  A default event is complete.
    class SYCLFFT_EXPORT event
    {
    public:
        event() = default;
        static event record(const queue &q); // record an event on q's stream now
        void wait() const;                   // hipEventSynchronize
        void wait_and_throw() const
        {
            wait();
        }
        [[nodiscard]] void *native_handle() const noexcept; // hipEvent_t or nullptr
        [[nodiscard]] bool complete() const;               // hipEventQuery == success

    private:
        std::shared_ptr<void> handle_;
        friend class event_access;
    };

    template <class Scalar>
    class SYCLFFT_EXPORT plan
    {
        static_assert(
            std::is_same_v<Scalar, float> || std::is_same_v<Scalar, double>,
            "syclfft supports float and double precision");

    public:
        using scalar_type = Scalar;
        using complex_type = complex<Scalar>;

        plan(const plan &) = delete;
        plan &operator=(const plan &) = delete;
        plan(plan &&) noexcept;
        plan &operator=(plan &&) noexcept;
        ~plan();

        event execute(complex_type *inout);
        event execute(complex_type *inout, const event &dependency);
        event execute(complex_type *inout, syclfft::span<const event> dependencies);

        event execute(const complex_type *input, complex_type *output);
        event execute(const complex_type *input, complex_type *output, const event &dependency);
        event execute(
            const complex_type *input, complex_type *output,
            syclfft::span<const event> dependencies);

        [[nodiscard]] syclfft::span<const std::size_t> shape() const noexcept;
        [[nodiscard]] std::size_t batch_count() const noexcept;
        [[nodiscard]] syclfft::direction direction() const noexcept;
        [[nodiscard]] syclfft::provider selected_provider() const noexcept;
        [[nodiscard]] std::size_t scratch_size_bytes() const noexcept;

        plan(
            queue q, std::vector<std::size_t> lengths, std::size_t batch_count,
            syclfft::direction direction, plan_options options);

    private:
        class impl;
        std::unique_ptr<impl> impl_;
    };

    SYCLFFT_EXPORT std::vector<provider_status> query_providers(const queue &q);

    template <class Scalar>
    plan<Scalar> plan_many_dft(
        queue q, syclfft::span<const std::size_t> lengths, std::size_t batch_count,
        syclfft::direction direction, plan_options options = {})
    {
        return plan<Scalar>(q, {lengths.begin(), lengths.end()}, batch_count, direction, options);
    }

    template <class Scalar>
    plan<Scalar> plan_many_dft(
        queue q, std::initializer_list<std::size_t> lengths, std::size_t batch_count,
        syclfft::direction direction, plan_options options = {})
    {
        return plan<Scalar>(q, {lengths.begin(), lengths.end()}, batch_count, direction, options);
    }

    template <class Scalar>
    plan<Scalar> plan_dft_1d(queue q, std::size_t n, syclfft::direction direction, plan_options options = {})
    {
        return plan_many_dft<Scalar>(q, {n}, 1, direction, options);
    }

    template <class Scalar>
    plan<Scalar> plan_dft_2d(
        queue q, std::size_t n0, std::size_t n1, syclfft::direction direction,
        plan_options options = {})
    {
        return plan_many_dft<Scalar>(q, {n0, n1}, 1, direction, options);
    }

    template <class Scalar>
    plan<Scalar> plan_dft_3d(
        queue q, std::size_t n0, std::size_t n1, std::size_t n2, syclfft::direction direction,
        plan_options options = {})
    {
        return plan_many_dft<Scalar>(q, {n0, n1, n2}, 1, direction, options);
    }

    extern template class plan<float>;
    extern template class plan<double>;

} // namespace syclfft
