// CPU emulation of the HIP device-side surface used by the provider kernels
// (TEST INFRASTRUCTURE ONLY; see hip_runtime_api.h).
#pragma once

#include "hip_runtime_api.h"

#include <atomic>
#include <cmath>
#include <cstdint>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define atomicAdd(p, v) emu_atomic_add_u32((p), (v))
#define __threadfence() std::atomic_thread_fence(std::memory_order_seq_cst)

#define __shared__ static

extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern thread_local dim3 blockDim;
extern thread_local dim3 gridDim;

void emu_syncthreads();
uint32_t emu_atomic_add_u32(uint32_t *p, uint32_t v);
// dynamic LDS of the running launch (sized by the launch's sharedMem argument)
unsigned char *emu_dynamic_lds();
#define __syncthreads() emu_syncthreads()

// Wave-level operations of the split-precision MFMA passes.  Every work-item of the
// workgroup calls them at the same program point (they synchronize the workgroup).
// Operand words hold two BF16 (prec 1) or FP16 (prec 2) values, element 0 in the low half.
uint32_t emu_to_bf16(float v);  // round to nearest even, bit pattern
float emu_from_bf16(uint32_t h);
uint32_t emu_to_f16(float v);   // round to nearest even (subnormals, overflow to inf)
float emu_from_f16(uint32_t h);
float emu_wave_xor(float v, int mask); // value of lane (lane ^ mask) of this wave of 64
double emu_wave_xor_d(double v, int mask);
// acc[j] += sum_k A[4(l/16)+j][k] B[k][l%16], the v_mfma_f32_16x16x32 lane layout
void emu_mfma_16x16x32(int prec, float *acc, const uint32_t *a, const uint32_t *b);

inline unsigned __umulhi(unsigned a, unsigned b)
{
    return static_cast<unsigned>((static_cast<unsigned long long>(a) * b) >> 32);
}

// run body() for every work-item of every workgroup of the grid
void emu_launch(dim3 grid, dim3 block, const std::function<void()> &body);

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                               \
    emu_launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })

// work-queue kernels take one sfh_queue_args argument: registered so that hipLaunchKernel
// calls them through `call` (every other entry takes one sfh_tile_args)
void emu_register_queue_kernel(const void *fn, void (*call)(void **args));

// tile kernels are launched through hipLaunchKernel with one sfh_tile_args argument
void emu_register_tile_kernel(const void *fn, void (*call)(const void *fn, void **args));

// agent-scope atomic accesses (the emulation runs workgroups on host threads)
#define __HIP_MEMORY_SCOPE_AGENT 3
template <typename T>
inline void __hip_atomic_store(T *p, T v, int, int)
{
    __atomic_store_n(p, v, __ATOMIC_SEQ_CST);
}
template <typename T>
inline T __hip_atomic_load(const T *p, int, int)
{
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}
template <typename T>
inline T __hip_atomic_fetch_add(T *p, T v, int, int)
{
    return emu_atomic_add_u32(p, v);
}
