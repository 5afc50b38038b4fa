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

// Real transforms of even length N = 2M run as complex transforms of length M over the
// interleaved pairs z[n] = x[2n] + i x[2n+1] (zero copy: the real buffer viewed as M complex
// values), plus one of these elementwise passes.  tw[k] = W_N^k = exp(-2 pi i k / N),
// k = 0..M, interleaved (re, im).
//
// R2C untangle, Z = DFT_M(z) -> X[k], k = 0..M (M + 1 outputs per transform):
//   X[k] = 1/2 (Z[k] + conj Z[M-k]) - i/2 W_N^k (Z[k] - conj Z[M-k]),  Z[M] = Z[0]
// (the even- and odd-sample spectra E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i
// combined as X[k] = E[k] + W_N^k O[k]), times `scale`.
template <typename T>
__global__ __launch_bounds__(256) void r2c_untangle(const T *z, T *x, const T *tw, uint64_t M, uint64_t count, T scale)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t b = i / (M + 1), k = i - b * (M + 1);
        const T *zb = z + 2 * b * M;
        const uint64_t k0 = k == M ? 0 : k, k1 = k == 0 ? 0 : M - k;
        const T zr = zb[2 * k0], zi = zb[2 * k0 + 1];
        const T mr = zb[2 * k1], mi = -zb[2 * k1 + 1]; // conj Z[M-k]
        const T ar = zr + mr, ai = zi + mi;             // Z[k] + conj Z[M-k]
        const T dr = zr - mr, di = zi - mi;             // Z[k] - conj Z[M-k]
        const T wr = tw[2 * k], wi = tw[2 * k + 1];
        const T tr = wr * dr - wi * di, ti = wr * di + wi * dr; // W d
        // -i (t) = (ti, -tr)
        x[2 * i] = T(0.5) * scale * (ar + ti);
        x[2 * i + 1] = T(0.5) * scale * (ai - tr);
    }
}

// C2R tangle, X[k], k = 0..M (Hermitian half spectrum) -> Z[k], k = 0..M-1, whose inverse
// DFT_M is x[2n] + i x[2n+1]:
//   Z[k] = (X[k] + conj X[M-k]) + i W_N^-k (X[k] - conj X[M-k])
// (the spectra of the even and odd samples, X[k] + X[k+M] and (X[k] - X[k+M]) W_N^-k, with
// X[k+M] = conj X[M-k]), times `scale`.
template <typename T>
__global__ __launch_bounds__(256) void c2r_tangle(const T *x, T *z, const T *tw, uint64_t M, uint64_t count, T scale)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t b = i / M, k = i - b * M;
        const T *xb = x + 2 * b * (M + 1);
        const T xr = xb[2 * k], xi = xb[2 * k + 1];
        const T mr = xb[2 * (M - k)], mi = -xb[2 * (M - k) + 1]; // conj X[M-k]
        const T ar = xr + mr, ai = xi + mi;
        const T dr = xr - mr, di = xi - mi;
        const T wr = tw[2 * k], wi = -tw[2 * k + 1]; // W_N^-k
        const T tr = wr * dr - wi * di, ti = wr * di + wi * dr;
        // + i t = (-ti, tr)
        z[2 * i] = scale * (ar - ti);
        z[2 * i + 1] = scale * (ai + tr);
    }
}
// Layout conversion of `batch` transforms of `n` complex values each (element x of transform
// b at b * dist + x * stride, in complex elements): interleaved (re, im) pairs at re[2 o],
// re[2 o + 1], or planar (split-complex) storage with the real parts at re[o] and the
// imaginary parts at im[o].  The grid's x dimension walks the elements of a transform and
// its y dimension the transforms, so no index division is needed.
template <typename T>
__global__ __launch_bounds__(256) void layout_copy(
    const T *s_re, const T *s_im, int64_t s_stride, int64_t s_dist, int s_planar, T *d_re, T *d_im,
    int64_t d_stride, int64_t d_dist, int d_planar, uint64_t n, uint64_t batch)
{
    const uint64_t xstep = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t b = blockIdx.y; b < batch; b += gridDim.y) {
        const int64_t sb = static_cast<int64_t>(b) * s_dist, db = static_cast<int64_t>(b) * d_dist;
        for (uint64_t x = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; x < n; x += xstep) {
            const int64_t so = sb + static_cast<int64_t>(x) * s_stride;
            const int64_t dop = db + static_cast<int64_t>(x) * d_stride;
            T vr, vi;
            if (s_planar) {
                vr = s_re[so];
                vi = s_im[so];
            } else {
                vr = s_re[2 * so];
                vi = s_re[2 * so + 1];
            }
            if (d_planar) {
                d_re[dop] = vr;
                d_im[dop] = vi;
            } else {
                d_re[2 * dop] = vr;
                d_re[2 * dop + 1] = vi;
            }
        }
    }
}
} // namespace

extern "C" int sfh_launch_real_pass(
    int c2r, const void *in, void *out, const void *tw, uint64_t M, uint64_t batch, int is_double, double scale,
    hipStream_t stream)
{
    const uint64_t count = batch * (c2r ? M : M + 1);
    if (count == 0) {
        return 0;
    }
    const uint64_t blocks64 = (count + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(blocks64 < 65536 ? blocks64 : 65536);
    if (is_double) {
        if (c2r) {
            hipLaunchKernelGGL(c2r_tangle<double>, dim3(blocks), dim3(256), 0, stream, static_cast<const double *>(in),
                               static_cast<double *>(out), static_cast<const double *>(tw), M, count, scale);
        } else {
            hipLaunchKernelGGL(r2c_untangle<double>, dim3(blocks), dim3(256), 0, stream,
                               static_cast<const double *>(in), static_cast<double *>(out),
                               static_cast<const double *>(tw), M, count, scale);
        }
    } else {
        const float sc = static_cast<float>(scale);
        if (c2r) {
            hipLaunchKernelGGL(c2r_tangle<float>, dim3(blocks), dim3(256), 0, stream, static_cast<const float *>(in),
                               static_cast<float *>(out), static_cast<const float *>(tw), M, count, sc);
        } else {
            hipLaunchKernelGGL(r2c_untangle<float>, dim3(blocks), dim3(256), 0, stream, static_cast<const float *>(in),
                               static_cast<float *>(out), static_cast<const float *>(tw), M, count, sc);
        }
    }
    return static_cast<int>(hipGetLastError());
}

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

extern "C" int sfh_launch_layout_copy(
    const void *s_re, const void *s_im, int64_t s_stride, int64_t s_dist, int s_planar, void *d_re, void *d_im,
    int64_t d_stride, int64_t d_dist, int d_planar, uint64_t n, uint64_t batch, int is_double, hipStream_t stream)
{
    if (n == 0 || batch == 0) {
        return 0;
    }
    const uint64_t bx = (n + 255) / 256;
    const dim3 grid(static_cast<unsigned>(bx < 4096 ? bx : 4096), static_cast<unsigned>(batch < 65535 ? batch : 65535));
    if (is_double) {
        hipLaunchKernelGGL(layout_copy<double>, grid, dim3(256), 0, stream, static_cast<const double *>(s_re),
                           static_cast<const double *>(s_im), s_stride, s_dist, s_planar, static_cast<double *>(d_re),
                           static_cast<double *>(d_im), d_stride, d_dist, d_planar, n, batch);
    } else {
        hipLaunchKernelGGL(layout_copy<float>, grid, dim3(256), 0, stream, static_cast<const float *>(s_re),
                           static_cast<const float *>(s_im), s_stride, s_dist, s_planar, static_cast<float *>(d_re),
                           static_cast<float *>(d_im), d_stride, d_dist, d_planar, n, batch);
    }
    return static_cast<int>(hipGetLastError());
}
