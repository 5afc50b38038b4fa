// Device plan API over the MI355X provider (semantics of the reference's tests/sycl_tests.cpp):
// 1-D cases of every path with the reference's inputs and tolerances, batched and multi-D
// cases, fp64, a round trip chained by an event, serialized repeated executions, two plans
// in flight, a dependency span, provider diagnostics and the error contract.
#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <iostream>
#include <syclfft/syclfft.hpp>
#include <type_traits>
#include <vector>

#include "test_utils.hpp"

namespace
{
    template <class T>
    struct dbuf
    {
        T *p{};
        std::size_t n{};
        explicit dbuf(std::size_t count)
          : n(count)
        {
            tu::check(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)) == hipSuccess, "hipMalloc");
        }
        ~dbuf()
        {
            (void)hipFree(p);
        }
        void write(const std::vector<T> &v)
        {
            tu::check(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess, "upload");
        }
        std::vector<T> read() const
        {
            std::vector<T> v(n);
            tu::check(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess, "download");
            return v;
        }
    };

    template <class S>
    std::vector<std::complex<S>> reference(std::vector<std::complex<S>> v, const std::vector<std::size_t> &shape,
                                           std::size_t batch, syclfft::direction dir, syclfft::normalization norm)
    {
        std::size_t tsize = 1;
        for (auto n : shape) {
            tsize *= n;
        }
        const long double sign = dir == syclfft::direction::forward ? -1.0L : 1.0L;
        const long double pi = 3.141592653589793238462643383279502884L;
        for (std::size_t axis = 0; axis < shape.size(); ++axis) {
            std::size_t stride = 1;
            for (std::size_t i = axis + 1; i < shape.size(); ++i) {
                stride *= shape[i];
            }
            const std::size_t n = shape[axis];
            auto out = v;
            for (std::size_t b = 0; b < batch; ++b) {
                for (std::size_t lin = 0; lin < tsize; ++lin) {
                    const std::size_t c = (lin / stride) % n, base = lin - c * stride;
                    std::complex<long double> acc{};
                    for (std::size_t k = 0; k < n; ++k) {
                        const long double a = sign * 2 * pi * static_cast<long double>((c * k) % n) / n;
                        const std::complex<S> x = v[b * tsize + base + k * stride];
                        acc += std::complex<long double>(x.real(), x.imag()) * std::complex<long double>(cosl(a), sinl(a));
                    }
                    out[b * tsize + lin] = {static_cast<S>(acc.real()), static_cast<S>(acc.imag())};
                }
            }
            v = out;
        }
        const S scale = static_cast<S>(syclfft::detail::normalization_scale(norm, dir, tsize));
        for (auto &x : v) {
            x *= scale;
        }
        return v;
    }

    template <class S>
    void run_1d(syclfft::queue &q, std::size_t n, syclfft::direction dir, syclfft::normalization norm,
                syclfft::placement placement)
    {
        std::vector<std::complex<S>> src(n);
        for (std::size_t i = 0; i < n; ++i) {
            src[i] = {static_cast<S>((i * 5 + 1) % 9) / S{4}, static_cast<S>((i * 3 + 2) % 7) / S{5}};
        }
        const auto expected = reference(src, {n}, 1, dir, norm);
        std::vector<syclfft::complex<S>> dsrc(src.begin(), src.end());
        dbuf<syclfft::complex<S>> in(n), out(n), rep(n);
        in.write(dsrc);
        auto fft = syclfft::plan_dft_1d<S>(q, n, dir, tu::opts().placed(placement).norm(norm).on(syclfft::provider::portable_sycl));
        const auto dep = syclfft::event::record(q);
        const bool inplace = placement == syclfft::placement::in_place;
        auto ev = inplace ? fft.execute(in.p, dep) : fft.execute(in.p, out.p, dep);
        syclfft::event repeated;
        if (!inplace) {
            repeated = fft.execute(in.p, rep.p); // second execution shares scratch: serialized
        }
        ev.wait_and_throw();
        repeated.wait_and_throw();
        const auto got = inplace ? in.read() : out.read();
        const S tol = std::is_same_v<S, float> ? S{4e-4} : S{2e-11};
        for (std::size_t i = 0; i < n; ++i) {
            const std::complex<S> a = got[i];
            tu::check(std::abs(a - expected[i]) <= tol * (S{1} + std::abs(expected[i])),
                      "1-D mismatch n=" + std::to_string(n) + " at " + std::to_string(i));
        }
        if (!inplace) {
            const auto r = rep.read();
            tu::check(std::memcmp(r.data(), got.data(), n * sizeof(r[0])) == 0, "repeated execution differs");
        }
        tu::check(fft.selected_provider() == syclfft::provider::hip_native, "provider");
    }

    template <class S>
    void run_many(syclfft::queue &q, std::vector<std::size_t> shape, std::size_t batch)
    {
        std::size_t count = batch;
        for (auto n : shape) {
            count *= n;
        }
        std::vector<std::complex<S>> src(count);
        for (std::size_t i = 0; i < count; ++i) {
            src[i] = {static_cast<S>((i * 11 + 2) % 17) / S{8}, static_cast<S>((i * 3 + 1) % 13) / S{7}};
        }
        const auto expected = reference(src, shape, batch, syclfft::direction::forward, syclfft::normalization::none);
        dbuf<syclfft::complex<S>> in(count), out(count);
        in.write(std::vector<syclfft::complex<S>>(src.begin(), src.end()));
        auto fft = syclfft::plan_many_dft<S>(q, shape, batch, syclfft::direction::forward);
        fft.execute(in.p, out.p).wait_and_throw();
        const auto got = out.read();
        const S tol = std::is_same_v<S, float> ? S{8e-4} : S{5e-11};
        for (std::size_t i = 0; i < count; ++i) {
            const std::complex<S> a = got[i];
            tu::check(std::abs(a - expected[i]) <= tol * (S{1} + std::abs(expected[i])), "many mismatch");
        }
    }
} // namespace

int main()
try {
    static_assert(sizeof(syclfft::complex<float>) == 2 * sizeof(float));
    static_assert(alignof(syclfft::complex<float>) == alignof(float));
    static_assert(std::is_trivially_copyable_v<syclfft::complex<float>>);
    static_assert(sizeof(syclfft::complex<double>) == sizeof(std::complex<double>));
    const std::complex<float> hv{2.5f, -1.25f};
    const syclfft::complex<float> dv{hv};
    tu::check(static_cast<std::complex<float>>(dv) == hv, "std::complex conversion");

    syclfft::queue q;
    using syclfft::direction;
    using syclfft::normalization;
    using syclfft::placement;
    for (auto norm : {normalization::none, normalization::forward, normalization::backward, normalization::orthogonal}) {
        run_1d<float>(q, 8, direction::forward, norm, placement::out_of_place);
    }
    run_1d<float>(q, 7, direction::backward, normalization::none, placement::in_place);
    run_1d<float>(q, 6, direction::forward, normalization::none, placement::out_of_place);
    run_1d<float>(q, 15, direction::backward, normalization::none, placement::out_of_place);
    run_1d<float>(q, 35, direction::forward, normalization::none, placement::out_of_place);
    run_1d<float>(q, 11, direction::backward, normalization::none, placement::in_place);
    run_1d<double>(q, 11, direction::forward, normalization::orthogonal, placement::out_of_place);
    run_1d<float>(q, 4096, direction::forward, normalization::none, placement::out_of_place);
    run_1d<double>(q, 2520, direction::backward, normalization::orthogonal, placement::out_of_place);
    run_many<float>(q, {3, 4}, 2);
    run_many<float>(q, {2, 3, 5}, 1);
    run_many<float>(q, {11, 3}, 1);
    run_many<double>(q, {16, 9, 10}, 2);

    { // round trip chained by an event (tests/sycl_tests.cpp:372-399)
        dbuf<syclfft::complex<float>> in(13), spec(13), back(13);
        std::vector<syclfft::complex<float>> src(13);
        for (std::size_t i = 0; i < 13; ++i) {
            src[i] = {static_cast<float>(i) / 7.0f, static_cast<float>(i % 3) / 5.0f};
        }
        in.write(src);
        auto f = syclfft::plan_dft_1d<float>(q, 13, direction::forward);
        auto b = syclfft::plan_dft_1d<float>(q, 13, direction::backward, tu::opts().norm(normalization::backward));
        const auto done = f.execute(in.p, spec.p);
        b.execute(spec.p, back.p, done).wait_and_throw();
        const auto r = back.read();
        for (std::size_t i = 0; i < 13; ++i) {
            tu::check(std::abs(static_cast<std::complex<float>>(r[i]) - static_cast<std::complex<float>>(src[i])) < 2e-3f, "round trip");
        }
    }
    { // provider diagnostics and automatic selection
        const auto ps = syclfft::query_providers(q);
        tu::check(!ps.empty() && ps.front().id == syclfft::provider::hip_native && ps.front().available, "hip_native unavailable");
        auto a = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        tu::check(a.selected_provider() == syclfft::provider::hip_native, "automatic selection");
    }
    { // two plans in flight, impulses (tests/sycl_tests.cpp:410-435)
        dbuf<syclfft::complex<float>> ia(8), ib(8), oa(8), ob(8);
        std::vector<syclfft::complex<float>> sa(8), sb(8);
        sa[0] = {1.0f, 0.0f};
        sb[1] = {1.0f, 0.0f};
        ia.write(sa);
        ib.write(sb);
        auto p1 = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        auto p2 = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        auto e1 = p1.execute(ia.p, oa.p);
        auto e2 = p2.execute(ib.p, ob.p);
        e1.wait_and_throw();
        e2.wait_and_throw();
        tu::check(std::abs(oa.read()[0].real() - 1.0f) < 1e-5f && std::abs(ob.read()[0].real() - 1.0f) < 1e-5f, "concurrent plans");
    }
    { // dependency span of two producers on another stream (tests/sycl_tests.cpp:483-506)
        hipStream_t s{};
        tu::check(hipStreamCreate(&s) == hipSuccess, "stream");
        syclfft::queue producer(q.device(), s);
        dbuf<syclfft::complex<float>> in(8), out(8);
        std::vector<syclfft::complex<float>> first(4), zero(4);
        first[0] = {1.0f, 0.0f};
        tu::check(hipMemcpyAsync(in.p, first.data(), 4 * sizeof(first[0]), hipMemcpyHostToDevice, s) == hipSuccess, "p1");
        const auto d1 = syclfft::event::record(producer);
        tu::check(hipMemcpyAsync(in.p + 4, zero.data(), 4 * sizeof(zero[0]), hipMemcpyHostToDevice, s) == hipSuccess, "p2");
        const auto d2 = syclfft::event::record(producer);
        const std::array<syclfft::event, 2> deps{d1, d2};
        auto fft = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        fft.execute(in.p, out.p, syclfft::span<const syclfft::event>{deps}).wait_and_throw();
        for (const auto &v : out.read()) {
            tu::check(std::abs(v.real() - 1.0f) < 1e-5f && std::abs(v.imag()) < 1e-5f, "dependency span");
        }
        (void)hipStreamDestroy(s);
    }
    using syclfft::error_code;
    tu::expect_error(error_code::provider_unavailable, [&] {
        (void)syclfft::plan_dft_1d<float>(q, 8, direction::forward, tu::opts().on(syclfft::provider::cufft));
    }, "cufft");
    tu::expect_error(error_code::invalid_argument, [&] { (void)syclfft::plan_dft_1d<float>(q, 8, static_cast<direction>(0)); }, "direction");
    tu::expect_error(error_code::invalid_pointer, [&] {
        auto f = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        f.execute(nullptr, nullptr);
    }, "null");
    tu::expect_error(error_code::invalid_pointer, [&] {
        auto f = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        std::vector<syclfft::complex<float>> host(8);
        dbuf<syclfft::complex<float>> out(8);
        f.execute(host.data(), out.p);
    }, "host pointer");
    tu::expect_error(error_code::invalid_pointer, [&] {
        auto f = syclfft::plan_dft_1d<float>(q, 8, direction::forward);
        dbuf<syclfft::complex<float>> small(4), out(8);
        f.execute(small.p, out.p);
    }, "buffer shorter than the transform");
    tu::expect_error(error_code::invalid_argument, [&] {
        auto f = syclfft::plan_dft_1d<float>(q, 8, direction::forward, tu::opts().in_place());
        dbuf<syclfft::complex<float>> a(8), b(8);
        f.execute(a.p, b.p);
    }, "placement");
    tu::expect_error(error_code::invalid_state, [&] {
        auto original = syclfft::plan_dft_1d<float>(q, 4, direction::forward);
        auto moved = std::move(original);
        (void)moved;
        original.execute(nullptr, nullptr); // NOLINT(bugprone-use-after-move)
    }, "moved-from");
    std::cout << "device tests passed\n";
    return 0;
} catch (const std::exception &ex) {
    std::cerr << "device tests failed: " << ex.what() << '\n';
    return 1;
}
