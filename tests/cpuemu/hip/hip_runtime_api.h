// CPU emulation of the small HIP surface the provider uses (TEST INFRASTRUCTURE ONLY).
// Kernels run on host threads: one std::thread per work-item of a workgroup, a
// std::barrier for __syncthreads, workgroups one after another.  "Device" memory is host
// heap memory, so AddressSanitizer checks every access the kernels make.
#pragma once

#include <cstddef>
#include <cstdint>

typedef int hipError_t;
enum
{
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorInvalidConfiguration = 9,
    hipErrorInvalidDevice = 101,
    hipErrorNotReady = 600
};
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
typedef void *hipDeviceptr_t;
enum hipMemcpyKind
{
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3
};
enum hipMemoryType
{
    hipMemoryTypeUnregistered = 0,
    hipMemoryTypeHost = 1,
    hipMemoryTypeDevice = 2,
    hipMemoryTypeManaged = 3
};
struct hipPointerAttribute_t
{
    hipMemoryType type;
    int device;
    void *devicePointer;
    void *hostPointer;
    int isManaged;
    unsigned allocationFlags;
};
struct hipDeviceProp_t
{
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
};
#define hipEventDisableTiming 2u
#define hipEventDisableSystemFence 0x20000000
enum hipFuncAttribute
{
    hipFuncAttributeMaxDynamicSharedMemorySize = 8
};

struct dim3
{
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1)
      : x(a)
      , y(b)
      , z(c)
    {}
};

hipError_t hipGetDeviceCount(int *);
hipError_t hipGetDevice(int *);
hipError_t hipSetDevice(int);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *, int);
hipError_t hipMalloc(void **, size_t);
hipError_t hipFree(void *);
hipError_t hipMemcpy(void *, const void *, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t);
#define hipHostMallocDefault 0u
hipError_t hipHostMalloc(void **, size_t, unsigned);
hipError_t hipHostFree(void *);
hipError_t hipMemsetAsync(void *, int, size_t, hipStream_t);
hipError_t hipMemset(void *, int, size_t);
hipError_t hipEventCreate(hipEvent_t *);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *, const void *);
hipError_t hipMemGetAddressRange(hipDeviceptr_t *, size_t *, hipDeviceptr_t);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipLaunchKernel(const void *, dim3, dim3, void **, size_t, hipStream_t);
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int);
enum hipDeviceAttribute_t
{
    hipDeviceAttributeMultiprocessorCount = 63
};
hipError_t hipDeviceGetAttribute(int *, hipDeviceAttribute_t, int);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *, const void *, int, size_t);
hipError_t hipEventQuery(hipEvent_t);
hipError_t hipStreamCreate(hipStream_t *);
typedef enum hipStreamCaptureStatus
{
    hipStreamCaptureStatusNone = 0,
    hipStreamCaptureStatusActive = 1
} hipStreamCaptureStatus;
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *);
#define hipStreamNonBlocking 1u
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
