// Small helper kernels of the provider.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace
{
// out[i] = scale * in[i] for interleaved complex values (the degenerate plan whose
// lengths are all 1; the reference's copy_scaled pass, src/sycl/sycl.cpp:87-98).
template <typename T>
__global__ __launch_bounds__(256) void copy_scale(const T *in, T *out, uint64_t n2, T scale)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += stride) {
        out[i] = in[i] * scale;
    }
}
} // namespace

extern "C" int sfh_launch_copy_scale(
    const void *in, void *out, uint64_t count, int is_double, double scale, hipStream_t stream)
{
    if (count == 0) {
        return 0;
    }
    const uint64_t n2 = 2 * count;
    const uint64_t blocks64 = (n2 + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(blocks64 < 65536 ? blocks64 : 65536);
    if (is_double) {
        hipLaunchKernelGGL(
            copy_scale<double>, dim3(blocks), dim3(256), 0, stream, static_cast<const double *>(in),
            static_cast<double *>(out), n2, scale);
    } else {
        hipLaunchKernelGGL(
            copy_scale<float>, dim3(blocks), dim3(256), 0, stream, static_cast<const float *>(in),
            static_cast<float *>(out), n2, static_cast<float>(scale));
    }
    return static_cast<int>(hipGetLastError());
}
