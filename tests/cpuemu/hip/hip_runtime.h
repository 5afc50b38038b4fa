// CPU emulation of the HIP device-side surface used by the provider kernels
// (TEST INFRASTRUCTURE ONLY; see hip_runtime_api.h).
#pragma once

#include "hip_runtime_api.h"

#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern thread_local dim3 blockDim;
extern thread_local dim3 gridDim;

void emu_syncthreads();
// dynamic LDS of the running launch (sized by the launch's sharedMem argument)
unsigned char *emu_dynamic_lds();
#define __syncthreads() emu_syncthreads()

inline unsigned __umulhi(unsigned a, unsigned b)
{
    return static_cast<unsigned>((static_cast<unsigned long long>(a) * b) >> 32);
}

// run body() for every work-item of every workgroup of the grid
void emu_launch(dim3 grid, dim3 block, const std::function<void()> &body);

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                               \
    emu_launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })

// tile kernels are launched through hipLaunchKernel with one sfh_tile_args argument
void emu_register_tile_kernel(const void *fn, void (*call)(const void *fn, void **args));
