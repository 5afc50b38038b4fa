/*
 * syclfft MI355X serving loop — C ABI (part of syclfft_provider_hip.so).
 *
 * A request loop for transform requests of mixed shapes that arrive over time.  Requests
 * carry device buffers; the server admits them while the HBM footprint of the work in
 * flight (request buffers plus the device memory of resident plans) stays within a budget
 * and while launch slots are free, merges waiting requests of the same shape into one
 * pointer-array execution (syclfft_hip_plan_execute_many), keeps plans (twiddles, chirp
 * spectra, scratch) resident and reuses them, and runs launches on several streams so that
 * requests admitted while others are on the GPU overlap with them.  Completion records give
 * each request's queueing time, run time and end-to-end latency.
 *
 * No reference counterpart: the reference has plans and events only, with one plan per
 * concurrent stream of work (README.md:45-47, tests/sycl_tests.cpp:410-435).  Statuses are
 * the provider's (SYCLFFT_HIP_STATUS_*, 1 + syclfft::error_code).
 */
#ifndef SYCLFFT_HIP_SERVER_H
#define SYCLFFT_HIP_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "syclfft_hip_provider.h"

#define SYCLFFT_HIP_SERVER_ABI_VERSION 1u

/* One transform request: `batch_count` transforms of `lengths` (C order, contiguous
 * batches) from device buffer `in` to device buffer `out` (in == out with in_place = 1). */
typedef struct syclfft_hip_request_v1
{
    uint32_t abi_version;   /* SYCLFFT_HIP_SERVER_ABI_VERSION */
    uint32_t scalar;        /* SYCLFFT_HIP_SCALAR_FLOAT / _DOUBLE */
    int32_t direction;      /* -1 forward, +1 backward */
    uint32_t rank;          /* 1..3 */
    uint64_t lengths[3];
    uint64_t batch_count;   /* >= 1 */
    uint32_t normalization; /* SYCLFFT_HIP_NORM_* */
    uint32_t in_place;
    const void *in;
    void *out;
    uint64_t tag;           /* caller's cookie, returned in the completion record */
} syclfft_hip_request_v1;

typedef struct syclfft_hip_completion_v1
{
    uint64_t ticket;
    uint64_t tag;
    int32_t status;          /* SYCLFFT_HIP_STATUS_* of the request */
    uint32_t group_size;     /* requests merged into the launch that ran this one */
    double queue_seconds;    /* submit -> launch enqueued */
    double run_seconds;      /* launch enqueued -> completion observed */
    double latency_seconds;  /* submit -> completion observed */
    double submit_at;        /* the three instants, in seconds since the server was created */
    double launch_at;
    double done_at;
} syclfft_hip_completion_v1;

typedef struct syclfft_hip_server_stats_v1
{
    uint64_t submitted;
    uint64_t completed;
    uint64_t failed;
    uint64_t launches;            /* pointer-array executions enqueued */
    uint64_t plans_created;
    uint64_t plans_resident;
    uint64_t plans_evicted;
    uint64_t peak_footprint_bytes;
    uint64_t budget_bytes;
} syclfft_hip_server_stats_v1;

/* Create a server on HIP device `device` (-1 = current): `memory_budget_bytes` of HBM for
 * request buffers in flight plus resident plans, `streams` launch streams (1..16), and at
 * most `max_batch_transforms` transforms merged into one launch (0 = 512). */
SYCLFFT_HIP_API void *syclfft_hip_server_create(
    int device, uint64_t memory_budget_bytes, uint32_t streams, uint64_t max_batch_transforms, char *err,
    size_t cap);

/* Queue a request; returns at once with its ticket.  Invalid requests (bad shape, host or
 * short buffers, a footprint above the budget) are rejected here with their status. */
SYCLFFT_HIP_API int syclfft_hip_server_submit(
    void *server, const syclfft_hip_request_v1 *request, uint64_t *ticket, char *err, size_t cap);

/* Block until the request has completed; fills *out and returns the request's status.
 * Each ticket can be waited for once. */
SYCLFFT_HIP_API int syclfft_hip_server_wait(
    void *server, uint64_t ticket, syclfft_hip_completion_v1 *out, char *err, size_t cap);

/* Block until every submitted request has completed. */
SYCLFFT_HIP_API int syclfft_hip_server_drain(void *server, char *err, size_t cap);

SYCLFFT_HIP_API void syclfft_hip_server_stats(void *server, syclfft_hip_server_stats_v1 *out);

/* Load replay for measurements: submit requests[k] when arrival_seconds[k] has elapsed since
 * the call started (busy-waiting on a steady clock; arrival_seconds = NULL submits them back
 * to back), or, with sequential != 0, each only after the previous one has completed; wait
 * for all of them and fill completions[k].  Submission and waiting run in native threads,
 * so the arrival process is not limited by the caller's language runtime.  Returns OK, or
 * the first failing request's status (its message in err). */
SYCLFFT_HIP_API int syclfft_hip_server_replay(
    void *server, const syclfft_hip_request_v1 *requests, const double *arrival_seconds, size_t count,
    int sequential, syclfft_hip_completion_v1 *completions, char *err, size_t cap);

/* Drain, then release streams and resident plans. */
SYCLFFT_HIP_API void syclfft_hip_server_destroy(void *server);

#endif /* SYCLFFT_HIP_SERVER_H */
