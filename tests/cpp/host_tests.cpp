// host::plan over the FFTW3-API host provider (semantics of the reference's
// tests/host_tests.cpp): every normalization, in/out of place, multi-D, round trip,
// and the error contract.  Expected values come from a brute-force DFT in long double.
#include <cmath>
#include <complex>
#include <iostream>
#include <limits>
#include <syclfft/host.hpp>
#include <vector>

#include "test_utils.hpp"

template <class S>
static std::vector<std::complex<S>> reference(std::vector<std::complex<S>> v, const std::vector<std::size_t> &shape,
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
static void compare(const std::vector<std::complex<S>> &a, const std::vector<std::complex<S>> &e, S tol, const char *what)
{
    for (std::size_t i = 0; i < a.size(); ++i) {
        tu::check(std::abs(a[i] - e[i]) <= tol * (S{1} + std::abs(e[i])), std::string(what) + " mismatch at " + std::to_string(i));
    }
}

template <class S>
static void run_case(std::vector<std::size_t> shape, std::size_t batch, syclfft::direction dir,
                     syclfft::normalization norm, syclfft::placement placement)
{
    std::size_t count = batch;
    for (auto n : shape) {
        count *= n;
    }
    std::vector<std::complex<S>> in(count);
    for (std::size_t i = 0; i < count; ++i) {
        in[i] = {static_cast<S>((i * 7) % 13) / S{7}, static_cast<S>((i * 5 + 3) % 11) / S{9}};
    }
    const auto expected = reference(in, shape, batch, dir, norm);
    auto fft = syclfft::host::plan_many_dft<S>(shape, batch, dir, tu::opts().placed(placement).norm(norm).on(syclfft::provider::fftw));
    tu::check(fft.selected_provider() == syclfft::provider::fftw, "wrong host provider");
    tu::check(fft.batch_count() == batch, "wrong batch count");
    tu::check(std::equal(fft.shape().begin(), fft.shape().end(), shape.begin()), "wrong shape");
    const S tol = std::is_same_v<S, float> ? S{2e-4} : S{1e-11};
    if (placement == syclfft::placement::in_place) {
        auto a = in;
        fft.execute(a.data());
        compare(a, expected, tol, "in-place");
    } else {
        std::vector<std::complex<S>> a(count);
        fft.execute(in.data(), a.data());
        compare(a, expected, tol, "out-of-place");
    }
}

int main()
try {
    for (auto norm : {syclfft::normalization::none, syclfft::normalization::forward, syclfft::normalization::backward,
                      syclfft::normalization::orthogonal}) {
        run_case<float>({8}, 2, syclfft::direction::forward, norm, syclfft::placement::out_of_place);
        run_case<double>({3, 5}, 1, syclfft::direction::backward, norm, syclfft::placement::in_place);
    }
    run_case<float>({2, 3, 4}, 2, syclfft::direction::backward, syclfft::normalization::none,
                    syclfft::placement::out_of_place);
    run_case<double>({13}, 3, syclfft::direction::forward, syclfft::normalization::orthogonal,
                     syclfft::placement::out_of_place);
    {
        std::vector<std::complex<double>> in(13), spec(13), back(13);
        for (std::size_t i = 0; i < 13; ++i) {
            in[i] = {static_cast<double>(i) / 7.0, static_cast<double>(i % 3) / 5.0};
        }
        auto f = syclfft::host::plan_dft_1d<double>(13, syclfft::direction::forward);
        auto b = syclfft::host::plan_dft_1d<double>(13, syclfft::direction::backward,
                                                     tu::opts().norm(syclfft::normalization::backward));
        f.execute(in.data(), spec.data());
        b.execute(spec.data(), back.data());
        compare(back, in, 2e-11, "round trip");
    }
    using syclfft::error_code;
    tu::expect_error(error_code::invalid_argument, [] { (void)syclfft::host::plan_dft_1d<float>(0, syclfft::direction::forward); }, "zero length");
    tu::expect_error(error_code::invalid_argument, [] {
        (void)syclfft::host::plan_many_dft<float>({std::numeric_limits<std::size_t>::max(), 2}, 1, syclfft::direction::forward);
    }, "overflow");
    tu::expect_error(error_code::invalid_argument, [] { (void)syclfft::host::plan_many_dft<float>({2, 2, 2, 2}, 1, syclfft::direction::forward); }, "rank 4");
    tu::expect_error(error_code::invalid_argument, [] { (void)syclfft::host::plan_dft_1d<float>(4, static_cast<syclfft::direction>(0)); }, "direction");
    tu::expect_error(error_code::invalid_pointer, [] {
        auto f = syclfft::host::plan_dft_1d<float>(4, syclfft::direction::forward);
        f.execute(nullptr, nullptr);
    }, "null pointers");
    tu::expect_error(error_code::invalid_argument, [] {
        auto f = syclfft::host::plan_dft_1d<float>(4, syclfft::direction::forward, tu::opts().in_place());
        std::vector<std::complex<float>> a(4), b(4);
        f.execute(a.data(), b.data());
    }, "placement");
    tu::expect_error(error_code::invalid_state, [] {
        auto original = syclfft::host::plan_dft_1d<float>(4, syclfft::direction::forward);
        auto moved = std::move(original);
        (void)moved;
        original.execute(nullptr, nullptr); // NOLINT(bugprone-use-after-move)
    }, "moved-from");
    std::cout << "host tests passed\n";
    return 0;
} catch (const std::exception &ex) {
    std::cerr << "host tests failed: " << ex.what() << '\n';
    return 1;
}
