// Minimal C++17 stand-in for std::span with the subset the plan API uses (dependency
// lists and shape views).  Same interface as the reference's include/syclfft/span.hpp.
#pragma once

#include <array>
#include <cstddef>
#include <type_traits>
#include <vector>

namespace syclfft
{

    template <class T>
    class span
    {
    public:
        using element_type = T;
        using value_type = std::remove_cv_t<T>;
        using size_type = std::size_t;
        using pointer = T *;
        using reference = T &;
        using iterator = T *;

        constexpr span() noexcept = default;
        constexpr span(T *data, size_type size) noexcept
          : data_(data)
          , size_(size)
        {}
        template <std::size_t N>
        constexpr span(T (&array)[N]) noexcept
          : data_(array)
          , size_(N)
        {}
        template <class U, std::size_t N, class = std::enable_if_t<std::is_convertible_v<U (*)[], T (*)[]>>>
        constexpr span(std::array<U, N> &array) noexcept
          : data_(array.data())
          , size_(N)
        {}
        template <class U, std::size_t N, class = std::enable_if_t<std::is_convertible_v<const U (*)[], T (*)[]>>>
        constexpr span(const std::array<U, N> &array) noexcept
          : data_(array.data())
          , size_(N)
        {}
        template <class U, class A, class = std::enable_if_t<std::is_convertible_v<U (*)[], T (*)[]>>>
        span(std::vector<U, A> &v) noexcept
          : data_(v.data())
          , size_(v.size())
        {}
        template <class U, class A, class = std::enable_if_t<std::is_convertible_v<const U (*)[], T (*)[]>>>
        span(const std::vector<U, A> &v) noexcept
          : data_(v.data())
          , size_(v.size())
        {}
        template <class U, class = std::enable_if_t<std::is_convertible_v<U (*)[], T (*)[]>>>
        constexpr span(const span<U> &other) noexcept
          : data_(other.data())
          , size_(other.size())
        {}

        [[nodiscard]] constexpr T *data() const noexcept
        {
            return data_;
        }
        [[nodiscard]] constexpr size_type size() const noexcept
        {
            return size_;
        }
        [[nodiscard]] constexpr bool empty() const noexcept
        {
            return size_ == 0;
        }
        constexpr T &operator[](size_type i) const noexcept
        {
            return data_[i];
        }
        constexpr iterator begin() const noexcept
        {
            return data_;
        }
        constexpr iterator end() const noexcept
        {
            return data_ + size_;
        }
        [[nodiscard]] constexpr span first(size_type count) const noexcept
        {
            return {data_, count};
        }
        [[nodiscard]] constexpr span last(size_type count) const noexcept
        {
            return {data_ + (size_ - count), count};
        }
        [[nodiscard]] constexpr span subspan(size_type offset) const noexcept
        {
            return {data_ + offset, size_ - offset};
        }
        [[nodiscard]] constexpr span subspan(size_type offset, size_type count) const noexcept
        {
            return {data_ + offset, count};
        }

    private:
        T *data_ = nullptr;
        size_type size_ = 0;
    };

} // namespace syclfft
