/*
 * syclfft MI355X-native device provider: pure C ABI.
 *
 * This is the drop-in boundary of the hot path.  It replaces the reference's device
 * provider ABI (include/syclfft/detail/cufft_provider_abi.hpp:14-47, a C++ ABI carrying
 * sycl::queue / sycl::event) and the built-in portable provider it stands beside
 * (src/sycl/sycl.cpp:405-799), with plain pointers, sizes and HIP handles only, so any
 * front-end (the C++ plan API in include/syclfft/syclfft.hpp, ctypes, cgo, JNI) can bind it.
 *
 * Streams and events are HIP handles passed as void*: `stream` is a hipStream_t (NULL =
 * the device's null stream), `deps` an array of hipEvent_t, `done` a hipEvent_t or NULL.
 *
 * Error convention (reference: include/syclfft/detail/provider_error_buffer.hpp:10-18,
 * src/providers/fftw/fftw_provider.cpp:310-355): functions never throw or abort; they
 * return NULL / a non-zero status and write a NUL-terminated message into err[cap]
 * (truncated to cap-1 bytes; err may be NULL or cap 0).  Status values are
 * SYCLFFT_HIP_STATUS_* below and map one-to-one onto syclfft::error_code.
 */
#ifndef SYCLFFT_HIP_PROVIDER_H
#define SYCLFFT_HIP_PROVIDER_H

#include <stddef.h>
#include <stdint.h>

#if defined(__cplusplus)
#define SYCLFFT_HIP_EXTERN_C extern "C"
#else
#define SYCLFFT_HIP_EXTERN_C
#endif
#define SYCLFFT_HIP_API SYCLFFT_HIP_EXTERN_C __attribute__((visibility("default")))

#define SYCLFFT_DEVICE_PROVIDER_ABI_VERSION 1u

/* scalar ids and direction values are those of include/syclfft/detail/provider_abi.h:23-29 */
#define SYCLFFT_HIP_SCALAR_FLOAT 1u
#define SYCLFFT_HIP_SCALAR_DOUBLE 2u

/* normalization: values of syclfft::normalization (include/syclfft/common.hpp:44-50) */
#define SYCLFFT_HIP_NORM_NONE 0u
#define SYCLFFT_HIP_NORM_FORWARD 1u
#define SYCLFFT_HIP_NORM_BACKWARD 2u
#define SYCLFFT_HIP_NORM_ORTHOGONAL 3u

/* fp32 arithmetic of the radix stages (no reference counterpart) */
#define SYCLFFT_HIP_PRECISION_VALU 0u       /* plain FP32/FP64 butterflies */
#define SYCLFFT_HIP_PRECISION_SPLIT_BF16 1u /* MFMA DFT stages, BF16 hi/lo x3, FP32 accumulate */
#define SYCLFFT_HIP_PRECISION_SPLIT_FP16 2u /* MFMA DFT stages, scaled FP16 hi/lo x3 */

/* status codes = 1 + syclfft::error_code (include/syclfft/common.hpp:83-93) */
#define SYCLFFT_HIP_STATUS_OK 0
#define SYCLFFT_HIP_STATUS_INVALID_ARGUMENT 1
#define SYCLFFT_HIP_STATUS_INVALID_STATE 2
#define SYCLFFT_HIP_STATUS_INVALID_POINTER 3
#define SYCLFFT_HIP_STATUS_UNSUPPORTED_PRECISION 4
#define SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE 5
#define SYCLFFT_HIP_STATUS_PLANNING_FAILED 6
#define SYCLFFT_HIP_STATUS_EXECUTION_FAILED 7
#define SYCLFFT_HIP_STATUS_PLUGIN_ERROR 8

/*
 * Plan configuration.  Fields abi_version..batch_count have the meaning of
 * cufft_plan_config_v2 (include/syclfft/detail/cufft_provider_abi.hpp:14-23) and of
 * syclfft_host_plan_config_v1 (include/syclfft/detail/provider_abi.h:31-41); lengths are
 * C order (last dimension contiguous), batches are contiguous with distance prod(lengths).
 * `normalization` is passed so the scale fuses into the last kernel instead of the
 * reference's trailing copy_scaled pass (src/sycl/sycl.cpp:769-773).
 */
typedef struct syclfft_hip_plan_config_v1
{
    uint32_t abi_version;   /* SYCLFFT_DEVICE_PROVIDER_ABI_VERSION */
    uint32_t scalar;        /* SYCLFFT_HIP_SCALAR_FLOAT / _DOUBLE */
    int32_t direction;      /* -1 forward, +1 backward */
    uint32_t rank;          /* 1..3 */
    uint64_t lengths[3];
    uint64_t batch_count;   /* >= 1 */
    uint32_t in_place;      /* 1: execute(inout) with in == out */
    uint32_t normalization; /* SYCLFFT_HIP_NORM_* */
    uint32_t measure;       /* 1: planning_mode::measure -> autotune tile shapes */
    uint32_t precision;     /* SYCLFFT_HIP_PRECISION_* (fp32 plans only) */
    int32_t device;         /* HIP device ordinal, -1 = current device */
    uint32_t reserved[3];   /* must be zero */
} syclfft_hip_plan_config_v1;

/* Function table, the shape of cufft_provider_v2 (cufft_provider_abi.hpp:25-34). */
typedef struct syclfft_device_provider_v1
{
    uint32_t abi_version;
    const char *name;
    int (*supports)(int device, char *err, size_t cap);
    void *(*create)(const syclfft_hip_plan_config_v1 *config, char *err, size_t cap);
    void (*destroy)(void *plan);
    uint64_t (*scratch_bytes)(const void *plan);
    int (*execute)(
        void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps,
        void *done, char *err, size_t cap);
} syclfft_device_provider_v1;

typedef const syclfft_device_provider_v1 *(*syclfft_get_device_provider_v1_fn)(void);

/* Getter resolved by the loader, as syclfft_get_cufft_provider_v2 (cufft_provider_abi.hpp:46-47)
 * and syclfft_get_host_provider_v1 (provider_abi.h:52) are. */
SYCLFFT_HIP_API const syclfft_device_provider_v1 *syclfft_get_device_provider_v1(void);

/* 1 if `device` (-1 = current) is a usable gfx950 device; else 0 and a reason.
 * Replaces cufft_provider_v2::supports and the USM-aspect check of the portable provider
 * (src/sycl/sycl.cpp:493-497, 912-919). */
SYCLFFT_HIP_API int syclfft_hip_supports(int device, char *err, size_t cap);

/* Build a plan: validates the config like checked_element_count
 * (src/core/provider_loader.cpp:112-132) and the direction check (src/sycl/sycl.cpp:435-438),
 * chooses the kernels, uploads FP64-generated twiddle/chirp tables and allocates scratch
 * (replaces impl::impl, src/sycl/sycl.cpp:424-591).  NULL on failure. */
SYCLFFT_HIP_API void *syclfft_hip_plan_create(
    const syclfft_hip_plan_config_v1 *config, char *err, size_t cap);

/* As syclfft_hip_plan_create, returning the status (SYCLFFT_HIP_STATUS_*) and the plan in
 * *plan (NULL on failure), so front-ends can map failures onto syclfft::error_code
 * (invalid_argument / unsupported_precision / provider_unavailable / planning_failed, as
 * src/sycl/sycl.cpp:424-591 throws them). */
SYCLFFT_HIP_API int syclfft_hip_plan_create_ex(
    const syclfft_hip_plan_config_v1 *config, void **plan, char *err, size_t cap);

/* Build a real plan of rank 1..3 whose last length N = lengths[rank - 1] is even (out of
 * place, precision VALU): direction -1 is R2C (batch_count x prod(lengths) real values in,
 * C order, and batch_count x prod(lengths) / N x (N/2 + 1) complex values out, the
 * non-redundant half of the spectrum along the last axis), +1 is C2R (the inverse, from a
 * Hermitian half spectrum).  Batches are contiguous with those distances; normalization
 * scales by prod(lengths) as for complex plans.  The last axis runs as a complex plan of
 * length N/2 over the real rows viewed as N/2 complex values plus one elementwise pass;
 * the leading axes run as strided complex axes over the half spectrum.  Executed, described
 * and destroyed like any other plan.  No reference counterpart: the reference plans C2C
 * transforms only (README.md:18-20). */
SYCLFFT_HIP_API int syclfft_hip_real_plan_create(
    const syclfft_hip_plan_config_v1 *config, void **plan, char *err, size_t cap);

/* Caller data layout of a layout plan, FFTW's advanced-interface strides plus split-complex
 * storage.  Element x (the row-major flat index over lengths[0..rank)) of transform b sits at
 * b * distance + x * stride, counted in complex elements for interleaved storage and in real
 * elements of each array for planar storage.  The reference's FFTW provider passes stride 1
 * and distance prod(lengths) (src/providers/fftw/fftw_provider.cpp:45-60, istride/idist and
 * ostride/odist of fftw_plan_many_dft); strides and split-complex storage are outside the
 * reference's scope (README.md:18-20). */
typedef struct syclfft_hip_layout_v1
{
    uint64_t in_stride;    /* >= 1 */
    uint64_t in_distance;  /* >= 1 */
    uint64_t out_stride;   /* >= 1 */
    uint64_t out_distance; /* >= 1 */
    uint32_t in_planar;    /* 1: input as separate real and imaginary arrays */
    uint32_t out_planar;   /* 1: output as separate real and imaginary arrays */
} syclfft_hip_layout_v1;

/* Build a C2C plan over a caller layout.  The config is that of syclfft_hip_plan_create_ex;
 * the layout must map distinct (transform, element) pairs to distinct elements (not
 * checked).  A rank-1 interleaved layout whose transform is one tile launch is addressed
 * directly by that launch (no extra memory traffic); every other layout runs a gather into
 * plan scratch, the contiguous plan in place there, and a scatter, i.e. two more passes
 * over the data.  Executed with syclfft_hip_layout_execute; syclfft_hip_plan_execute also
 * accepts it when both sides are interleaved.  No reference counterpart (README.md:18-20). */
SYCLFFT_HIP_API int syclfft_hip_layout_plan_create(
    const syclfft_hip_plan_config_v1 *config, const syclfft_hip_layout_v1 *layout, void **plan, char *err,
    size_t cap);

/* Execute a layout plan: `in` / `out` are the interleaved arrays, or the real-part arrays of
 * planar sides, whose imaginary-part arrays are in_imag / out_imag (NULL for interleaved
 * sides).  In-place plans take in == out and in_imag == out_imag; out-of-place plans take
 * non-overlapping input and output ranges.  Ordering, dependencies and `done` as for
 * syclfft_hip_plan_execute. */
SYCLFFT_HIP_API int syclfft_hip_layout_execute(
    void *plan, const void *in, const void *in_imag, void *out, void *out_imag, void *stream, void *const *deps,
    size_t ndeps, void *done, char *err, size_t cap);

/* Free tables and scratch after the plan's last execution has finished
 * (replaces impl::~impl, src/sycl/sycl.cpp:593-613). */
SYCLFFT_HIP_API void syclfft_hip_plan_destroy(void *plan);

/* Device scratch owned by the plan, in bytes (plan<Scalar>::scratch_size_bytes,
 * src/sycl/sycl.cpp:901-907). */
SYCLFFT_HIP_API uint64_t syclfft_hip_plan_scratch_bytes(const void *plan);

/* Enqueue the transform on `stream` after every event in deps[0..ndeps) and after the
 * plan's previous execution (executions of one plan serialize because they share scratch,
 * src/sycl/sycl.cpp:637-641); record `done` when it is complete.  Asynchronous: returns
 * once the kernels are enqueued.  in/out are device pointers to batch*prod(lengths)
 * interleaved complex values; in == out is required for in-place plans and rejected for
 * out-of-place plans (src/sycl/sycl.cpp:619-635).  Replaces impl::execute
 * (src/sycl/sycl.cpp:615-776) and cufft_provider_v2::execute.
 * May be captured into a HIP graph (kernel nodes only); under capture the ordering after
 * the plan's previous execution is left to the captured stream order, and the plan must
 * outlive the graph. */
SYCLFFT_HIP_API int syclfft_hip_plan_execute(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps,
    void *done, char *err, size_t cap);

/* As syclfft_hip_plan_execute for complex plans, but allocating nothing: launches that pass
 * intermediates through scratch (four-step axes, the three-launch Bluestein form) use a
 * file-scope device array of syclfft_hip_noalloc_scratch_limit() bytes (64 MiB) per device,
 * reserved when the library is loaded on a device and held until the process ends, instead
 * of the plan's own scratch.  A plan whose scratch_bytes exceed that limit, and real or
 * layout plans, are rejected with SYCLFFT_HIP_STATUS_INVALID_ARGUMENT.  Every call on one
 * device shares that device's array: calls whose executions overlap on one device race,
 * so the caller orders them.  Results are bit-identical to syclfft_hip_plan_execute.
 * No counterpart in the reference: for callers that cannot allocate device memory while
 * they run (HIP graph capture, serving loops, harnesses that forbid hipMalloc). */
SYCLFFT_HIP_API int syclfft_hip_plan_execute_noalloc(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps,
    void *done, char *err, size_t cap);

/* Bytes of the device array syclfft_hip_plan_execute_noalloc uses as scratch. */
SYCLFFT_HIP_API uint64_t syclfft_hip_noalloc_scratch_limit(void);

/* Enqueue `count` transforms of the plan's shape (1 <= count <= the plan's batch count)
 * given by host arrays of device pointers: transform k reads ins[k] and writes outs[k]
 * (ins[k] == outs[k] for in-place plans).  Same ordering, dependency and completion
 * semantics as syclfft_hip_plan_execute; the pointer arrays are copied before the call
 * returns.  One plan then serves any number of independent buffers up to its batch count
 * in one set of launches (the serving loop batches requests this way).  No reference
 * counterpart: the reference executes one contiguous batch per call
 * (src/sycl/sycl.cpp:615-776).  Under HIP graph capture only contiguous pointers
 * (ins[k] = ins[0] + k * prod(lengths) elements, likewise outs) are accepted; others
 * need a table upload and return INVALID_STATE. */
SYCLFFT_HIP_API int syclfft_hip_plan_execute_many(
    void *plan, uint64_t count, const void *const *ins, void *const *outs, void *stream, void *const *deps,
    size_t ndeps, void *done, char *err, size_t cap);

/* Enqueue `count` executions of possibly different plans on `stream`: execution k runs
 * plans[k] over its whole batch from ins[k] into outs[k], with the checks and ordering of
 * syclfft_hip_plan_execute; `done` (optional) is recorded when all are complete.  Plans
 * whose execution is one plain tile launch (7-smooth lengths that fit one tile) and share
 * a kernel variant run together in ONE persistent launch: its workgroups take tiles from a
 * device work queue, so transforms of uneven lengths, batch counts, directions and
 * normalizations share a dispatch.  Other plans execute first, as by
 * syclfft_hip_plan_execute.  All plans must belong to one device.  The work list is
 * uploaded from host staging, so this call cannot be captured into a HIP graph
 * (INVALID_STATE).  No reference counterpart: the reference runs one plan per execute
 * (src/sycl/sycl.cpp:615-776); this is the batching entry for callers that hold many small
 * plans, such as a request loop. */
SYCLFFT_HIP_API int syclfft_hip_execute_queue(
    void *const *plans, const void *const *ins, void *const *outs, size_t count, void *stream, void *done,
    char *err, size_t cap);

/* Human-readable kernel sequence of a plan (one line per launch) for tracing and tests;
 * returns the full length (like snprintf). No reference counterpart. */
SYCLFFT_HIP_API size_t syclfft_hip_plan_describe(const void *plan, char *buf, size_t cap);

/* Number of kernel launches one execute enqueues. No reference counterpart. */
SYCLFFT_HIP_API uint32_t syclfft_hip_plan_launch_count(const void *plan);

#endif /* SYCLFFT_HIP_PROVIDER_H */
