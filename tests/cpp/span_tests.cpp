// syclfft::span behaviour (semantics of the reference's tests/span_tests.cpp).
#include <array>
#include <iostream>
#include <syclfft/span.hpp>
#include <vector>

#include "test_utils.hpp"

static int sum(syclfft::span<const int> v)
{
    int s = 0;
    for (int x : v) {
        s += x;
    }
    return s;
}

int main()
try {
    int values[]{1, 2, 3, 4};
    syclfft::span<int> m{values};
    syclfft::span<const int> c{m};
    tu::check(c.data() == values && c.size() == 4, "array span");
    tu::check(sum(syclfft::span<int>{values}) == 10, "converting span");
    const std::array<int, 3> a{{5, 6, 7}};
    syclfft::span<const int> as{a};
    tu::check(as.size() == 3 && as[2] == 7, "std::array span");
    const std::vector<int> v{8, 9, 10};
    syclfft::span<const int> vs{v};
    tu::check(vs.first(2)[1] == 9 && vs.last(2)[0] == 9 && vs.subspan(1)[1] == 10, "views");
    const syclfft::span<const int> e;
    tu::check(e.empty() && e.size() == 0 && e.data() == nullptr && e.begin() == e.end(), "empty span");
    tu::check(e.last(0).empty() && e.subspan(0).empty(), "empty views");
    std::cout << "span tests passed\n";
    return 0;
} catch (const std::exception &ex) {
    std::cerr << "span tests failed: " << ex.what() << '\n';
    return 1;
}
