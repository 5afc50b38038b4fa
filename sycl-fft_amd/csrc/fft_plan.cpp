// Plan compiler and executor of the MI355X provider (C ABI: include/syclfft_hip_provider.h).
//
// Reference behaviour reproduced here:
//   * validation: checked_element_count (src/core/provider_loader.cpp:112-132) and the
//     direction check (src/sycl/sycl.cpp:435-438);
//   * axis order 0..rank-1, stride = product of trailing lengths (src/sycl/sycl.cpp:63-70,
//     692-767);
//   * algorithm per axis: power of two / 7-smooth / otherwise Bluestein, decided by trial
//     division by 2, 3, 5, 7 (supported_factors, src/sycl/sycl.cpp:48-61, 703, 715);
//   * normalization scale (normalization_scale, src/core/provider_loader.cpp:134-147);
//   * serialized executions of one plan (src/sycl/sycl.cpp:637-641, 774).
// Differences by design: every axis runs in 1 launch (fits LDS), 2 launches (four-step
// split) or 3 launches (Bluestein with a convolution longer than LDS); the scale and the
// direction are fused into the first/last launch; Bluestein uses the smallest 7-smooth
// convolution length >= 2n-1 instead of the next power of two.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

#include "fft_args.h"
#include "fft_kernels.h"
#include "fft_tables.h"
#include "syclfft_hip_provider.h"

namespace sfh
{

// ------------------------------------------------------------------------------------
// Errors
// ------------------------------------------------------------------------------------

struct sfh_error : std::runtime_error
{
    int status;
    sfh_error(int st, const std::string &msg)
      : std::runtime_error(msg)
      , status(st)
    {}
};

static void set_err(char *err, size_t cap, const std::string &msg)
{
    if (!err || cap == 0) {
        return;
    }
    const size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = '\0';
}

static void hip_check(hipError_t st, const char *what)
{
    if (st != hipSuccess) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_EXECUTION_FAILED,
            std::string(what) + ": " + hipGetErrorString(st));
    }
}

static void trace(const std::string &msg)
{
    const char *enabled = std::getenv("SYCLFFT_TRACE");
    if (!enabled || enabled[0] == '\0' || enabled[0] == '0') {
        return;
    }
    std::fprintf(stderr, "[syclfft] hip provider: %s\n", msg.c_str());
    std::fflush(stderr);
}

// ------------------------------------------------------------------------------------
// Integer helpers
// ------------------------------------------------------------------------------------

sfh_fastdiv make_fastdiv(uint32_t d)
{
    if (d == 0 || d >= (1u << 31)) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "divisor out of range");
    }
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d) {
        ++s;
    }
    const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
    return {d, static_cast<uint32_t>(m), s};
}

static bool is_pow2(uint64_t v)
{
    return v && !(v & (v - 1));
}

static uint32_t ilog2(uint64_t v)
{
    uint32_t r = 0;
    while ((uint64_t(1) << r) < v) {
        ++r;
    }
    return r;
}

// supported_factors (src/sycl/sycl.cpp:48-61): 7-smooth test by trial division.
static bool is_smooth7(uint64_t n)
{
    for (uint64_t r : {2u, 3u, 5u, 7u}) {
        while (n % r == 0) {
            n /= r;
        }
    }
    return n == 1;
}

// Radix program for a 7-smooth length: radix-16 passes for the power-of-two part (radix 8
// when the kernel variant stages more than 16 elements per lane), then 8/4/2 for the
// remainder, then 9s, 3s, 5s and 7s.
static std::vector<uint32_t> radix_plan(uint64_t n, uint32_t maxr)
{
    std::vector<uint32_t> r;
    uint32_t a = 0;
    while (n % 2 == 0) {
        n /= 2;
        ++a;
    }
    while (maxr >= 16 && a >= 4) {
        r.push_back(16);
        a -= 4;
    }
    while (a >= 3) {
        r.push_back(8);
        a -= 3;
    }
    if (a == 3) {
        r.push_back(8);
    } else if (a == 2) {
        r.push_back(4);
    } else if (a == 1) {
        r.push_back(2);
    }
    while (n % 9 == 0) {
        r.push_back(9);
        n /= 9;
    }
    for (uint32_t p : {3u, 5u, 7u}) {
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    }
    if (n != 1) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "length is not 7-smooth");
    }
    return r;
}

// Smallest 2^a 3^b 5^c 7^d >= x satisfying pred.
template <class Pred>
static uint64_t smooth_at_least(uint64_t x, Pred pred)
{
    uint64_t best = 0;
    for (uint64_t p7 = 1; p7 < 4 * x; p7 *= 7) {
        for (uint64_t p5 = p7; p5 < 4 * x; p5 *= 5) {
            for (uint64_t p3 = p5; p3 < 4 * x; p3 *= 3) {
                uint64_t v = p3;
                while (v < x) {
                    v *= 2;
                }
                if ((best == 0 || v < best) && pred(v)) {
                    best = v;
                }
            }
        }
    }
    return best;
}

// ------------------------------------------------------------------------------------
// Plan
// ------------------------------------------------------------------------------------

enum buf_id
{
    BUF_IN = 0,
    BUF_OUT = 1,
    BUF_SCRATCH = 2
};

struct launch
{
    int variant{};
    int mode{};
    sfh_tile_args args{};
    buf_id src{BUF_IN}, dst{BUF_OUT};
    int64_t src_off{0}, dst_off{0}; // element offsets into the source / destination buffer
    uint32_t grid{};                // workgroups launched
    // table offsets into the plan's table block (patched into args at plan time)
    int64_t tw_off{-1}, pre_off{-1}, mid_off{-1}, post_off{-1};
    // the store conjugates before POST (Bluestein's inverse FFT, conj-FFT-conj): folded into
    // the epilogue sign with a conjugated post table
    bool conj_post{false};
    // transform dimension of the tile grid (SFH_UNIT_*), its extent at the plan's batch
    // count, and the element strides between its units on each side
    uint32_t unit_dim{0};
    uint64_t unit_extent{0};
    int64_t unit_in_stride{0}, unit_out_stride{0};
    // length-specialized kernel replacing the generic entry (null: generic), with its
    // per-pass twiddle tables (offset into the plan's table block, then device address)
    const sfh_fixed_kernel *fixed{nullptr};
    // no generic variant holds this launch's tile: only `fixed` can run it
    bool fixed_only{false};
    int64_t fx_tw_off{-1};
    const void *fx_twtab{nullptr};
    std::string desc;
};

// Planner choices the autotuner searches over (planning_mode::measure).
struct tuning
{
    int split{0};    // 1: multi-launch form (four-step, Bluestein A/B/C) even where one launch fits
    int wshift{0};   // column-tile width scaled by 2^wshift (-1, 0, +1)
    int prefetch{0}; // persistent prefetching form of plain / four-step-A launches
    int grow{0};     // the next larger kernel variant (more work-items per tile)

    std::string str() const
    {
        return "s" + std::to_string(split) + "w" + std::to_string(wshift) + "p" + std::to_string(prefetch) + "g"
               + std::to_string(grow);
    }
    static bool parse(const std::string &text, tuning &t)
    {
        int s = 0, w = 0, pf = 0, g = 0;
        if (std::sscanf(text.c_str(), "s%dw%dp%dg%d", &s, &w, &pf, &g) != 4) {
            return false;
        }
        if (s < 0 || s > 1 || w < -1 || w > 1 || pf < 0 || pf > 1 || g < 0 || g > 1) {
            return false;
        }
        t.split = s;
        t.wshift = w;
        t.prefetch = pf;
        t.grow = g;
        return true;
    }
};

// One staging slot of the work-queue ring (see execute_queue): device block, pinned host
// copy, and the event after which the slot's last launch has completed; `gen` counts its
// launches, so a plan can tell whether the slot still holds the launch it was part of.
struct queue_slot
{
    void *dev{};
    void *host{};
    size_t bytes{};
    hipEvent_t free{};
    bool used{};
    std::atomic<uint64_t> gen{0};
};

struct plan
{
    syclfft_hip_plan_config_v1 cfg{};
    tuning tune{};
    int device{};
    bool dbl{};
    size_t esize{};
    uint64_t total{};     // batch * prod(lengths)
    uint64_t transform{}; // prod(lengths)
    double scale{1.0};
    std::vector<launch> launches;
    std::vector<unsigned char> host_tables;
    std::map<std::string, int64_t> table_cache;
    void *tables{};
    uint64_t tables_bytes{};
    uint64_t scratch_elems{};
    void *scratch{};
    // Bluestein pair launches: one arrival counter per transform (zeroed at creation)
    uint64_t pair_transforms{};
    uint32_t *pair_counts{};
    hipEvent_t last{};
    bool have_last{};
    hipStream_t last_stream{}; // stream of the last execution
    // the last execution was part of a work-queue launch: its completion is that slot's event
    // while the slot's generation is still last_qgen (the slot is reused only after it completes)
    queue_slot *last_queue{};
    uint64_t last_qgen{};
    std::mutex mu;
    const sfh_kernel_variant *variants{};
    int nvariants{};
    int copy_only{}; // all lengths 1: a scaled copy
    int cap_limit{};  // largest tile (complex elements) the planner may use; 0 = no limit
    int split_launches{}; // launches whose radix-16 passes run on MFMA (split precision)
    // real 1-D plans (syclfft_hip_real_plan_create): 1 = R2C, 2 = C2R of even length N = 2 real_m,
    // run as the complex plan `half` of length real_m over the real buffer viewed as complex,
    // plus one untangle / tangle pass with the table W_N^k (k = 0..real_m) into / out of scratch
    int real{};
    struct plan *half{};
    // 1-D C2R plan whose half-length plan is one length-specialized launch: launches[0] runs
    // the tangle fused into that launch's first pass (one HBM pass)
    bool real_fused{};
    uint64_t real_m{};
    void *real_tw{};
    void *real_scratch{};
    // layout plans (syclfft_hip_layout_plan_create): the caller's strided / planar layout over
    // the contiguous plan `core`.  lay_fused: core's one launch addresses the caller's buffers
    // with the layout's strides; otherwise gather into lay_scratch, core in place there, scatter.
    int layout{};
    syclfft_hip_layout_v1 lay{};
    struct plan *core{};
    bool lay_fused{};
    void *lay_scratch{};
    // unit-table ring for executions over pointer arrays (allocated on first use): per slot
    // a device table block, its pinned host staging copy, and the event after which the
    // slot's last execution no longer reads it
    static constexpr int k_unit_slots = 4;
    uint64_t unit_bytes{};
    void *unit_dev[k_unit_slots]{};
    void *unit_host[k_unit_slots]{};
    hipEvent_t unit_free[k_unit_slots]{};
    bool unit_used[k_unit_slots]{};
    int unit_next{};
};

// Completion event of the plan's last execution (null: known complete).
static hipEvent_t last_event(const plan &p)
{
    if (!p.have_last) {
        return nullptr;
    }
    if (p.last_queue) {
        return p.last_queue->gen.load() == p.last_qgen ? p.last_queue->free : nullptr;
    }
    return p.last;
}

// Order `stream` after the plan's last execution (nothing to do on the same stream).
static void wait_last(const plan &p, hipStream_t stream)
{
    if (!p.have_last || stream == p.last_stream) {
        return;
    }
    if (hipEvent_t e = last_event(p)) {
        if (hipStreamWaitEvent(stream, e, 0) != hipSuccess) {
            (void)hipGetLastError();
            throw std::runtime_error("hipStreamWaitEvent(last) failed");
        }
    }
}

// Block the host until the plan's last execution has completed.
static void sync_last(const plan &p)
{
    if (hipEvent_t e = last_event(p)) {
        (void)hipEventSynchronize(e);
    }
}

static void set_last(plan &p, hipStream_t stream)
{
    p.have_last = true;
    p.last_stream = stream;
    p.last_queue = nullptr;
}

static bool variant_allowed(const plan &p, int i)
{
    if (p.variants[i].tw_global) {
        return false; // global-table twins replace their LDS-table variant after planning
    }
    return p.cap_limit <= 0 || p.variants[i].nt * p.variants[i].e <= p.cap_limit;
}

// The global-table twin of variant i (same threads and elements per thread), or -1.
static int global_table_twin(const plan &p, int i)
{
    if (i < 0 || p.variants[i].tw_global) {
        return -1;
    }
    for (int j = 0; j < p.nvariants; ++j) {
        if (p.variants[j].tw_global && p.variants[j].nt == p.variants[i].nt && p.variants[j].e == p.variants[i].e) {
            return j;
        }
    }
    return -1;
}

static int max_cap(const plan &p)
{
    int best = 0;
    for (int i = 0; i < p.nvariants; ++i) {
        if (variant_allowed(p, i)) {
            best = std::max(best, p.variants[i].nt * p.variants[i].e);
        }
    }
    return best;
}

// Sizes of the LDS twiddle tables of a launch: a two-level table for the pass twiddles of
// length L (2^tw_bits + ceil(L / 2^tw_bits) entries, the split minimizing the count) and,
// when the launch applies four-step twiddles of length N, a three-level table
// (2 * 2^t4_bits + ceil(N / 2^(2 t4_bits)) entries).
struct tw_layout
{
    uint32_t tw_bits{}, tw_entries{}, t4_bits{}, t4_entries{};
};

static tw_layout twiddle_layout(uint64_t L, uint64_t N)
{
    tw_layout t;
    uint64_t best = UINT64_MAX;
    for (uint32_t b = 0; b <= 16; ++b) {
        const uint64_t K = uint64_t(1) << b;
        const uint64_t e = K + (L + K - 1) / K;
        if (e < best) {
            best = e;
            t.tw_bits = b;
        }
    }
    t.tw_entries = static_cast<uint32_t>(best);
    if (N) {
        uint32_t b = (ilog2(N) + 2) / 3;
        b = std::max<uint32_t>(b, 1);
        const uint64_t K = uint64_t(1) << b;
        t.t4_bits = b;
        t.t4_entries = static_cast<uint32_t>(2 * K + (N + K * K - 1) / (K * K));
    }
    return t;
}

// LDS pitch of a tile column: L slots, plus one pad slot per 32 elements when the padded
// layout is used, plus one slot when several columns share the tile so that lanes on
// different columns at the same element hit different banks.
static uint32_t lds_pitch(uint32_t L, uint32_t W, bool pad)
{
    const uint32_t slots = pad ? L + (L + 31u) / 32u : L;
    return W > 1 ? slots + 1 : slots;
}

static uint64_t lds_bytes_of(const plan &p, uint32_t L, uint32_t W, const tw_layout &t, uint32_t *tab_off,
                             bool pad = false)
{
    const uint64_t pitch = lds_pitch(L, W, pad);
    const uint64_t tab = (pitch * W + 1) & ~uint64_t(1);
    if (tab_off) {
        *tab_off = static_cast<uint32_t>(tab);
    }
    return (tab + t.tw_entries + t.t4_entries) * p.esize;
}

static constexpr uint64_t k_lds_limit = 163840; // 160 KiB of LDS per workgroup on gfx950

// Smallest variant holding W columns of length L (at the LDS pitch) plus the tables.
static int pick_variant(const plan &p, uint32_t L, uint32_t W, const tw_layout &t)
{
    int best = -1;
    for (int i = 0; i < p.nvariants; ++i) {
        if (!variant_allowed(p, i)) {
            continue;
        }
        const int cap = p.variants[i].nt * p.variants[i].e;
        if (static_cast<uint64_t>(L) * W > static_cast<uint64_t>(cap)) {
            continue;
        }
        if (lds_bytes_of(p, L, W, t, nullptr) > k_lds_limit) {
            continue;
        }
        if (static_cast<uint32_t>(p.variants[i].nt) < W) {
            continue;
        }
        if (W > 1 && !is_pow2(static_cast<uint64_t>(p.variants[i].nt))) {
            continue;
        }
        if (best < 0 || cap < p.variants[best].nt * p.variants[best].e) {
            best = i;
        }
    }
    if (best >= 0 && p.tune.grow) {
        // the next larger qualifying variant, if any
        int next = -1;
        const int bcap = p.variants[best].nt * p.variants[best].e;
        for (int i = 0; i < p.nvariants; ++i) {
            const int cap = p.variants[i].nt * p.variants[i].e;
            if (!variant_allowed(p, i) || cap <= bcap || lds_bytes_of(p, L, W, t, nullptr) > k_lds_limit
                || static_cast<uint32_t>(p.variants[i].nt) < W
                || (W > 1 && !is_pow2(static_cast<uint64_t>(p.variants[i].nt)))) {
                continue;
            }
            if (next < 0 || cap < p.variants[next].nt * p.variants[next].e) {
                next = i;
            }
        }
        if (next >= 0) {
            best = next;
        }
    }
    return best;
}

static int pick_variant(const plan &p, uint32_t L, uint32_t W)
{
    return pick_variant(p, L, W, twiddle_layout(L, 0));
}

template <class T>
static int64_t add_table(plan &p, const std::string &key, const std::vector<cxl> &values)
{
    const auto it = p.table_cache.find(key);
    if (it != p.table_cache.end()) {
        return it->second;
    }
    size_t off = (p.host_tables.size() + 255) & ~size_t(255);
    p.host_tables.resize(off + values.size() * 2 * sizeof(T));
    T *dst = reinterpret_cast<T *>(p.host_tables.data() + off);
    for (size_t i = 0; i < values.size(); ++i) {
        dst[2 * i] = static_cast<T>(values[i].re);
        dst[2 * i + 1] = static_cast<T>(values[i].im);
    }
    p.table_cache[key] = static_cast<int64_t>(off);
    return static_cast<int64_t>(off);
}

static int64_t table(plan &p, const std::string &key, const std::vector<cxl> &values)
{
    return p.dbl ? add_table<double>(p, key, values) : add_table<float>(p, key, values);
}

// The table block a launch stages into LDS (layout of sfh_tile_args::twtab).
static int64_t twiddle_tables(plan &p, uint64_t L, uint64_t N, const tw_layout &t)
{
    const std::string key = "tw" + std::to_string(L) + "_" + std::to_string(N);
    if (p.table_cache.count(key)) {
        return p.table_cache[key];
    }
    std::vector<cxl> v;
    v.reserve(t.tw_entries + t.t4_entries);
    const uint64_t K = uint64_t(1) << t.tw_bits;
    for (uint64_t x = 0; x < K; ++x) {
        v.push_back(root_ld(x, L));
    }
    for (uint64_t y = 0; y < t.tw_entries - K; ++y) {
        v.push_back(root_ld(y * K, L));
    }
    if (N) {
        const uint64_t K4 = uint64_t(1) << t.t4_bits;
        for (uint64_t x = 0; x < K4; ++x) {
            v.push_back(root_ld(x, N));
        }
        for (uint64_t x = 0; x < K4; ++x) {
            v.push_back(root_ld(x * K4, N));
        }
        for (uint64_t y = 0; y < t.t4_entries - 2 * K4; ++y) {
            v.push_back(root_ld(y * K4 * K4, N));
        }
    }
    return table(p, key, v);
}

static void set_passes(launch &l, uint32_t L, bool second_fft, uint32_t maxr)
{
    const std::vector<uint32_t> radices = radix_plan(L, maxr);
    const uint32_t np = static_cast<uint32_t>(radices.size());
    if (np == 0 || (second_fft ? 2 * np : np) > SFH_MAX_PASSES) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "unsupported radix program");
    }
    auto fill = [&](uint32_t base, const std::vector<uint32_t> &rs) {
        uint32_t ns = 1;
        for (uint32_t i = 0; i < rs.size(); ++i) {
            sfh_pass &ps = l.args.pass[base + i];
            ps.radix = rs[i];
            ps.ns = ns;
            ps.twstride = L / (ns * rs[i]);
            ps.ns_div = make_fastdiv(ns);
            ns *= rs[i];
        }
    };
    fill(0, radices);
    // per-radix pass counts in the kernel's fixed order; radix_plan emits that order
    static const uint32_t order[8] = {16, 8, 4, 2, 9, 3, 5, 7};
    uint32_t pos = 0;
    uint32_t counts[8] = {};
    for (int k = 0; k < 8; ++k) {
        while (pos < radices.size() && radices[pos] == order[k]) {
            ++counts[k];
            ++pos;
        }
    }
    if (pos != radices.size()) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "radix program out of kernel order");
    }
    // FFT #1 without pass 0 (the first pass of the first non-empty group); FFT #2 without its
    // first pass (the last radix of FFT #1, i.e. the last non-empty group)
    for (int k = 0; k < 8; ++k) {
        l.args.rcount[k] = static_cast<uint8_t>(counts[k]);
        l.args.rcount2[k] = static_cast<uint8_t>(counts[k]);
    }
    for (int k = 0; k < 8; ++k) {
        if (counts[k]) {
            --l.args.rcount[k];
            break;
        }
    }
    for (int k = 7; k >= 0; --k) {
        if (counts[k]) {
            --l.args.rcount2[k];
            break;
        }
    }
    if (second_fft && np < 2) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "Bluestein length needs two radix passes");
    }
    if (second_fft) {
        std::vector<uint32_t> rev(radices.rbegin(), radices.rend());
        fill(np, rev);
    }
    l.args.npass = np;
    l.args.L = L;
    l.args.L_div = make_fastdiv(L);
}

struct geometry
{
    uint32_t L{}, W{};
    uint64_t wcount{}, nblo{1}, nbhi{1};
    int64_t in_w{}, in_j{}, in_blo{}, in_bhi{};
    int64_t out_w{}, out_k{}, out_blo{}, out_bhi{};
    uint64_t col_div{1}, col_mod{1}, col_bmul{0};
    uint32_t unit_dim{0}; // SFH_UNIT_*: the dimension that enumerates the axis' outer lines
};

static void check_span(const geometry &g)
{
    const auto span = [&](int64_t sw, int64_t sj) {
        return std::llabs(sw) * (g.W - 1) + std::llabs(sj) * (g.L - 1);
    };
    if (span(g.in_w, g.in_j) >= (int64_t(1) << 31) || span(g.out_w, g.out_k) >= (int64_t(1) << 31)
        || g.wcount >= (uint64_t(1) << 31) || g.col_div >= (uint64_t(1) << 31)
        || g.col_mod >= (uint64_t(1) << 31)) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_PLANNING_FAILED, "transform geometry exceeds 32-bit tile offsets");
    }
}

// LDS bank model: cycles one wave's access takes when lane l touches `esize` bytes at
// element slot slot[l] (-1 = inactive), with 64 banks of 4 bytes; lanes reading the same
// dword share it.  Ideal is esize / 4 cycles for 64 active lanes.
static uint32_t lds_cycles(const int64_t *slot, uint32_t esize)
{
    uint32_t words[64 * 4];
    uint32_t n = 0;
    for (int l = 0; l < 64; ++l) {
        if (slot[l] < 0) {
            continue;
        }
        for (uint32_t d = 0; d < esize / 4; ++d) {
            words[n++] = static_cast<uint32_t>(slot[l] * (esize / 4) + d);
        }
    }
    std::sort(words, words + n);
    n = static_cast<uint32_t>(std::unique(words, words + n) - words);
    uint32_t per_bank[64] = {};
    uint32_t worst = 0;
    for (uint32_t k = 0; k < n; ++k) {
        worst = std::max(worst, ++per_bank[words[k] % 64]);
    }
    return worst;
}

// Bank-model cost (cycles of the first wave) of the LDS traffic across one pass boundary:
// the writer's Stockham stores (radix wr at ns, or FFT #2's first pass at ns = 1 with
// writer_mid) and the next pass's reads (radix rr; 0 = no LDS reader), with the tile in
// layout `pad` at pitch LP.
static uint64_t boundary_cost(const sfh_tile_args &a, uint32_t esize, uint32_t wr, uint32_t wns, bool writer_mid,
                              uint32_t rr, bool pad, uint32_t LP)
{
    const uint32_t L = a.L, W = a.W;
    uint32_t lane_w[64], lane_i[64];
    for (uint32_t l = 0; l < 64; ++l) {
        if (W == 1) {
            lane_w[l] = 0;
            lane_i[l] = l;
        } else if (a.flags & SFH_WFAST) {
            lane_w[l] = l & (W - 1);
            lane_i[l] = l >> a.logW;
        } else {
            lane_w[l] = l >> a.logtpc;
            lane_i[l] = l & (a.tpc - 1);
        }
    }
    const auto slot = [pad](uint64_t e) { return static_cast<int64_t>(pad ? e + (e >> 5) : e); };
    uint64_t cost = 0;
    int64_t addr[64];
    const uint32_t mw = L / wr;
    for (uint32_t t = 0; t < wr; ++t) {
        for (uint32_t l = 0; l < 64; ++l) {
            const uint32_t i = lane_i[l];
            if (i >= mw) {
                addr[l] = -1;
                continue;
            }
            uint64_t e;
            if (writer_mid) {
                e = static_cast<uint64_t>(i) * wr + t;
            } else {
                const uint64_t kq = i / wns;
                e = kq * wns * wr + (i - kq * wns) + static_cast<uint64_t>(t) * wns;
            }
            addr[l] = static_cast<int64_t>(lane_w[l]) * LP + slot(e);
        }
        cost += lds_cycles(addr, esize);
    }
    if (rr) {
        const uint32_t mr = L / rr;
        for (uint32_t t = 0; t < rr; ++t) {
            for (uint32_t l = 0; l < 64; ++l) {
                const uint32_t i = lane_i[l];
                addr[l] = i < mr ? static_cast<int64_t>(lane_w[l]) * LP + slot(i + static_cast<uint64_t>(t) * mr) : -1;
            }
            cost += lds_cycles(addr, esize);
        }
    }
    return cost;
}

// LDS layout of every tile state of the launch's program, per pass boundary: layout 0 or
// the padded layout, by the bank model of the boundary's writes and reads.  The padded
// layout is eligible where both sides stay linear in t (writer: power-of-two radix and
// ns; reader: m = L/R a multiple of 32 or a power of two, or the per-element STORE_LDS
// loop).  Program step
// s runs pass[s] for FFT #1 and pass[s + 1] for FFT #2, whose pass 0 is fused into step
// np - 1 (which therefore writes with FFT #2's first radix at ns = 1).  Two candidates
// are compared: every boundary in layout 0 at the plain pitch, and the padded pitch with
// each boundary in its cheaper layout; pads are dropped when the padded tile does not fit.
static void set_lds_layout(const plan &p, launch &l, const tw_layout &tl, bool second_fft)
{
    sfh_tile_args &a = l.args;
    const uint32_t np = a.npass;
    const uint32_t total = second_fft ? 2 * np - 1 : np;
    const bool store_lds = (a.flags & SFH_STORE_LDS) != 0;
    const auto pidx = [np](uint32_t s) { return s < np ? s : s + 1; };
    const uint32_t es = static_cast<uint32_t>(p.esize);
    const uint32_t lp0 = lds_pitch(a.L, a.W, false), lp1 = lds_pitch(a.L, a.W, true);
    const bool pad_fits = lds_bytes_of(p, a.L, a.W, tl, nullptr, true) <= k_lds_limit;
    uint64_t cost0 = 0, cost1 = 0;
    std::vector<uint32_t> choice(total, 0);
    for (uint32_t s = 0; s < total; ++s) {
        const bool last = s + 1 == total;
        if (last && !store_lds) {
            continue; // the last pass writes HBM
        }
        const sfh_pass &ps = a.pass[pidx(s)];
        const bool mid = second_fft && s + 1 == np;
        const uint32_t wr = mid ? a.pass[np].radix : ps.radix;
        const uint32_t wns = mid ? 1u : ps.ns;
        const uint32_t rr = last ? 0u : a.pass[pidx(s + 1)].radix;
        const uint64_t c0 = boundary_cost(a, es, wr, wns, mid, rr, false, lp0);
        cost0 += c0;
        const bool eligible =
            is_pow2(wr) && is_pow2(wns) && (last || (a.L / rr) % 32 == 0 || is_pow2(a.L / rr));
        const uint64_t c1 = boundary_cost(a, es, wr, wns, mid, rr, false, lp1);
        const uint64_t c1p = eligible ? boundary_cost(a, es, wr, wns, mid, rr, true, lp1) : c1;
        choice[s] = c1p < c1 ? 1u : 0u;
        cost1 += std::min(c1, c1p);
    }
    const bool use_pad = pad_fits && cost1 < cost0;
    uint32_t prev = 0; // pass 0 reads HBM, or LDS filled in layout 0 by the load loop
    for (uint32_t s = 0; s < total; ++s) {
        sfh_pass &ps = a.pass[pidx(s)];
        ps.pad_in = prev;
        ps.pad_out = use_pad ? choice[s] : 0u;
        prev = ps.pad_out;
    }
    a.pad_store = store_lds ? prev : 0u;
    a.LP = use_pad ? lp1 : lp0;
    a.lds_bytes = static_cast<uint32_t>(lds_bytes_of(p, a.L, a.W, tl, &a.lds_tab, use_pad));
}

static launch make_launch(plan &p, int mode, geometry g, buf_id src, buf_id dst,
                          bool second_fft, const std::string &what, uint64_t t4n = 0)
{
    launch l;
    l.mode = mode;
    l.src = src;
    l.dst = dst;
    sfh_tile_args &a = l.args;
    const tw_layout tl = twiddle_layout(g.L, t4n);
    l.variant = pick_variant(p, g.L, g.W, tl);
    while (l.variant < 0 && g.W > 1) {
        g.W /= 2; // the tables did not fit beside a tile this wide
        l.variant = pick_variant(p, g.L, g.W, tl);
    }
    if (l.variant < 0) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "no kernel variant fits the tile");
    }
    check_span(g);
    const sfh_kernel_variant &v = p.variants[l.variant];
    set_passes(l, g.L, second_fft, static_cast<uint32_t>(v.maxr));
    a.W = g.W;
    a.logW = ilog2(g.W);
    a.tpc = static_cast<uint32_t>(v.nt) / g.W;
    a.logtpc = ilog2(a.tpc);
    a.wcount = g.wcount;
    a.nwt = (g.wcount + g.W - 1) / g.W;
    a.nblo = g.nblo;
    a.ntiles = a.nwt * g.nblo * g.nbhi;
    a.in_w = static_cast<int32_t>(g.in_w);
    a.in_j = static_cast<int32_t>(g.in_j);
    a.out_w = static_cast<int32_t>(g.out_w);
    a.out_k = static_cast<int32_t>(g.out_k);
    a.in_blo = g.in_blo;
    a.in_bhi = g.in_bhi;
    a.out_blo = g.out_blo;
    a.out_bhi = g.out_bhi;
    a.col_div = make_fastdiv(static_cast<uint32_t>(g.col_div));
    a.col_mod = make_fastdiv(static_cast<uint32_t>(g.col_mod));
    a.col_bmul = static_cast<uint32_t>(g.col_bmul);
    a.out_sx = 1.0;
    a.out_sy = 1.0;
    // lane mapping: consecutive lanes on consecutive columns when the source is
    // column-contiguous; otherwise along the column.
    const bool in_wc = (g.in_w == 1 && g.W > 1);
    const bool out_wc = (g.out_w == 1 && g.W > 1);
    if (in_wc) {
        a.flags |= SFH_WFAST;
        if (!out_wc) {
            a.flags |= SFH_STORE_LDS;
        }
    } else if (out_wc) {
        a.flags |= SFH_STORE_LDS;
    }
    a.tw_bits = tl.tw_bits;
    a.tw_entries = tl.tw_entries;
    a.t4_bits = tl.t4_bits;
    a.t4_entries = tl.t4_entries;
    set_lds_layout(p, l, tl, second_fft);
    l.unit_dim = g.unit_dim;
    if (g.unit_dim == SFH_UNIT_COL) {
        l.unit_extent = g.wcount;
        l.unit_in_stride = g.in_w;
        l.unit_out_stride = g.out_w;
    } else if (g.unit_dim == SFH_UNIT_BLO) {
        l.unit_extent = g.nblo;
        l.unit_in_stride = g.in_blo;
        l.unit_out_stride = g.out_blo;
    } else if (g.unit_dim == SFH_UNIT_BHI) {
        l.unit_extent = g.nbhi;
        l.unit_in_stride = g.in_bhi;
        l.unit_out_stride = g.out_bhi;
    }
    a.unit_dim = g.unit_dim;
    l.tw_off = twiddle_tables(p, g.L, t4n, tl);
    l.grid = static_cast<uint32_t>(std::min<uint64_t>(a.ntiles, 0xffffffffu));
    char buf[256];
    std::snprintf(
        buf, sizeof(buf), "%s L=%u W=%u tiles=%llu nt=%d e=%d mode=%d flags=0x%x lds=%u lp=%u", what.c_str(),
        g.L, g.W, static_cast<unsigned long long>(a.ntiles), v.nt, v.e, mode, a.flags, a.lds_bytes, a.LP);
    l.desc = buf;
    return l;
}

static const sfh_fixed_kernel *find_fixed(const plan &p, int mode_index, uint64_t L);
static const sfh_fixed_kernel *find_fixed_cols(const plan &p, int mode_index, uint64_t n, uint64_t wcount);
static int64_t fixed_tables(plan &p, const sfh_fixed_kernel &k);

// A launch that only a length-specialized kernel runs (no generic variant holds its tile):
// one contiguous transform of k.L points per workgroup, columns `g.wcount`.
static launch make_fixed_launch(plan &p, const sfh_fixed_kernel &k, const geometry &g, buf_id src, buf_id dst,
                                const std::string &what)
{
    launch l;
    l.mode = SFH_MI_PLAIN;
    l.src = src;
    l.dst = dst;
    l.variant = 0; // unused: the fixed kernel runs every execution of this launch
    l.fixed_only = true;
    sfh_tile_args &a = l.args;
    a.L = static_cast<uint32_t>(g.L);
    a.LP = a.L;
    a.W = 1;
    a.logW = 0;
    a.tpc = static_cast<uint32_t>(k.nt);
    a.wcount = g.wcount;
    a.nwt = g.wcount;
    a.nblo = 1;
    a.ntiles = g.wcount;
    a.in_w = static_cast<int32_t>(g.in_w);
    a.in_j = 1;
    a.out_w = static_cast<int32_t>(g.out_w);
    a.out_k = 1;
    a.col_div = make_fastdiv(1);
    a.col_mod = make_fastdiv(1);
    a.out_sx = 1.0;
    a.out_sy = 1.0;
    a.unit_dim = SFH_UNIT_COL;
    l.unit_dim = SFH_UNIT_COL;
    l.unit_extent = g.wcount;
    l.unit_in_stride = g.in_w;
    l.unit_out_stride = g.out_w;
    l.grid = static_cast<uint32_t>(std::min<uint64_t>(a.ntiles, 0xffffffffu));
    l.fixed = &k;
    l.fx_tw_off = fixed_tables(p, k);
    a.lds_bytes = k.lds_bytes;
    l.desc = what + " L=" + std::to_string(g.L) + " W=1 tiles=" + std::to_string(a.ntiles) + " fixed="
             + std::to_string(k.nt) + "x" + std::to_string(k.L);
    return l;
}

// One launch of a column-tile length-specialized kernel: W adjacent columns (of g.wcount)
// per tile, lanes column-fastest, each column's points at stride g.in_j / g.out_k, rows of a
// tile at g.in_w / g.out_w (plain mode).
static launch make_fixed_cols_launch(plan &p, const sfh_fixed_kernel &k, geometry g, buf_id src, buf_id dst,
                                     const std::string &what)
{
    g.L = k.L;
    g.W = static_cast<uint32_t>(k.W);
    check_span(g);
    launch l;
    l.mode = SFH_MI_PLAIN;
    l.src = src;
    l.dst = dst;
    l.variant = 0; // unused: the fixed kernel runs every execution of this launch
    l.fixed_only = true;
    sfh_tile_args &a = l.args;
    a.L = g.L;
    a.LP = a.L;
    a.W = g.W;
    a.logW = ilog2(g.W);
    a.tpc = static_cast<uint32_t>(k.nt) / g.W;
    a.logtpc = ilog2(a.tpc);
    a.wcount = g.wcount;
    a.nwt = (g.wcount + g.W - 1) / g.W; // the last tile of a row may hold fewer columns
    a.nblo = g.nblo;
    a.ntiles = a.nwt * g.nblo * g.nbhi;
    a.in_w = static_cast<int32_t>(g.in_w);
    a.in_j = static_cast<int32_t>(g.in_j);
    a.out_w = static_cast<int32_t>(g.out_w);
    a.out_k = static_cast<int32_t>(g.out_k);
    a.in_blo = g.in_blo;
    a.in_bhi = g.in_bhi;
    a.out_blo = g.out_blo;
    a.out_bhi = g.out_bhi;
    a.col_div = make_fastdiv(static_cast<uint32_t>(g.col_div));
    a.col_mod = make_fastdiv(static_cast<uint32_t>(g.col_mod));
    a.col_bmul = static_cast<uint32_t>(g.col_bmul);
    a.out_sx = 1.0;
    a.out_sy = 1.0;
    a.flags |= SFH_WFAST;
    // groups of 2^xcd_lg adjacent tiles on one XCD (SYCLFFT_HIP_XCD_PAIR=0: off;
    // SYCLFFT_HIP_XCD_GROUP=<lg>: groups of 2^lg tiles, lg = 1 .. 4, for measurements)
    const char *xp = std::getenv("SYCLFFT_HIP_XCD_PAIR");
    if (!(xp && std::strcmp(xp, "0") == 0)) {
        uint32_t lg = static_cast<uint32_t>(std::min(4, std::max(0, k.xcd_lg)));
        if (const char *xg = std::getenv("SYCLFFT_HIP_XCD_GROUP")) {
            lg = static_cast<uint32_t>(std::min(4, std::max(1, std::atoi(xg))));
        }
        a.flags |= lg << SFH_TILE_XCD_SHIFT;
    }
    l.unit_dim = g.unit_dim;
    if (g.unit_dim == SFH_UNIT_BLO) {
        l.unit_extent = g.nblo;
        l.unit_in_stride = g.in_blo;
        l.unit_out_stride = g.out_blo;
    } else if (g.unit_dim == SFH_UNIT_BHI) {
        l.unit_extent = g.nbhi;
        l.unit_in_stride = g.in_bhi;
        l.unit_out_stride = g.out_bhi;
    }
    a.unit_dim = g.unit_dim;
    l.grid = static_cast<uint32_t>(std::min<uint64_t>(a.ntiles, 0xffffffffu));
    l.fixed = &k;
    l.fx_tw_off = fixed_tables(p, k);
    a.lds_bytes = k.lds_bytes;
    l.desc = what + " L=" + std::to_string(g.L) + " W=" + std::to_string(g.W) + " tiles=" + std::to_string(a.ntiles)
             + " fixed=" + std::to_string(k.nt) + "x" + std::to_string(k.L) + " cols";
    return l;
}

// Columns per tile for an axis whose columns are `contig` (adjacent in memory) or whose
// points are contiguous.
static uint32_t choose_w(const plan &p, uint64_t L, uint64_t wcount, bool contig_cols)
{
    const uint64_t cap = static_cast<uint64_t>(max_cap(p));
    uint64_t w;
    if (contig_cols) {
        w = 16;
        while (w < 64 && w * 2 * L <= 2048) {
            w *= 2;
        }
    } else {
        w = 1;
        while (w < 64 && w * L < 1024) {
            w *= 2;
        }
    }
    if (p.tune.wshift > 0 && w < 64) {
        w *= 2;
    } else if (p.tune.wshift < 0 && w > 1) {
        w /= 2;
    }
    while (w > 1 && w * (L + 1) > cap) {
        w /= 2;
    }
    // do not make tiles wider than the column dimension needs
    while (w > 1 && w / 2 >= wcount) {
        w /= 2;
    }
    return static_cast<uint32_t>(w);
}

static bool fits_single(const plan &p, uint64_t L, uint32_t W)
{
    return L <= static_cast<uint64_t>(max_cap(p)) && pick_variant(p, static_cast<uint32_t>(L), W) >= 0;
}

// Split n = n1 * n2 (n1 <= n2) with both sides fitting a tile of `wmin` columns.
static bool split_two(const plan &p, uint64_t n, uint32_t wmin, uint64_t &n1, uint64_t &n2)
{
    uint64_t best = 0;
    for (uint64_t d = 2; d * d <= n; ++d) {
        if (n % d == 0) {
            const uint64_t e = n / d;
            if (fits_single(p, d, wmin) && fits_single(p, e, wmin)) {
                best = d; // largest d <= sqrt(n) wins
            }
        }
    }
    if (!best) {
        return false;
    }
    n1 = best;
    n2 = n / best;
    return true;
}

static uint32_t min_cols(const plan &p)
{
    return p.dbl ? 8u : 16u;
}

// One axis of extent n with `outer` lines before and `inner` elements after it.
static void plan_fourstep(plan &p, uint64_t outer, uint64_t n, uint64_t inner, uint64_t n1, uint64_t n2,
                          buf_id src, buf_id dst, int64_t off);
static void plan_bluestein(plan &p, uint64_t outer, uint64_t n, uint64_t inner, buf_id src, buf_id dst);

// Lines of an axis processed per launch group when the axis needs scratch between
// launches: with SYCLFFT_HIP_CHUNK_MB > 0 the group's intermediate (that many MiB) can stay
// in the 256 MiB Infinity Cache between the launches that write and read it.  Off by
// default (one launch group per axis).
static uint64_t chunk_lines(const plan &p, uint64_t outer, uint64_t line_elems)
{
    uint64_t mb = 0; // measured: the extra launch boundaries cost more than the cache reuse gains
    if (const char *e = std::getenv("SYCLFFT_HIP_CHUNK_MB")) {
        mb = std::strtoull(e, nullptr, 10);
    }
    if (mb == 0) {
        return outer;
    }
    const uint64_t bytes = mb << 20;
    const uint64_t per = std::max<uint64_t>(1, bytes / std::max<uint64_t>(1, line_elems * p.esize));
    return std::min(per, outer);
}

static void plan_axis(plan &p, uint64_t outer, uint64_t n, uint64_t inner, buf_id src, buf_id dst)
{
    if (n == 1) {
        return;
    }
    if (is_smooth7(n)) {
        // ---------------- one launch of a length-specialized kernel beyond the generic tiles
        if (inner == 1 && !p.tune.split && n > static_cast<uint64_t>(max_cap(p)) && n <= 0xffffffffu) {
            if (const sfh_fixed_kernel *k = find_fixed(p, SFH_MI_PLAIN, n)) {
                geometry g;
                g.L = static_cast<uint32_t>(n);
                g.W = 1;
                g.wcount = outer;
                g.in_w = g.out_w = static_cast<int64_t>(n);
                p.launches.push_back(make_fixed_launch(p, *k, g, src, dst, "single"));
                return;
            }
        }
        // ---------------- one launch of a column-tile length-specialized kernel (strided axis)
        // (tile offsets are 32-bit: the axis' span must stay below 2^31 elements)
        if (inner > 1 && !p.tune.split && inner * (n - 1) + 64 < (uint64_t(1) << 31)) {
            if (const sfh_fixed_kernel *k = find_fixed_cols(p, SFH_MI_PLAIN, n, inner)) {
                geometry g;
                g.wcount = inner;
                g.nblo = outer;
                g.unit_dim = SFH_UNIT_BLO;
                g.in_w = g.out_w = 1;
                g.in_j = g.out_k = static_cast<int64_t>(inner);
                g.in_blo = g.out_blo = static_cast<int64_t>(n * inner);
                p.launches.push_back(make_fixed_cols_launch(p, *k, g, src, dst, "single"));
                return;
            }
        }
        // ---------------- single launch ----------------
        {
            geometry g;
            g.L = static_cast<uint32_t>(std::min<uint64_t>(n, 0xffffffffu));
            if (inner == 1) {
                g.W = choose_w(p, n, outer, false);
                g.wcount = outer;
                g.unit_dim = SFH_UNIT_COL;
                g.in_w = g.out_w = static_cast<int64_t>(n);
                g.in_j = g.out_k = 1;
            } else {
                g.W = choose_w(p, n, inner, true);
                g.wcount = inner;
                g.nblo = outer;
                g.unit_dim = SFH_UNIT_BLO;
                g.in_w = g.out_w = 1;
                g.in_j = g.out_k = static_cast<int64_t>(inner);
                g.in_blo = g.out_blo = static_cast<int64_t>(n * inner);
            }
            uint64_t s1 = 0, s2 = 0;
            bool can_split = p.tune.split && split_two(p, n, 1, s1, s2);
            if (inner > 1 && !can_split && n <= static_cast<uint64_t>(max_cap(p))) {
                // A strided axis whose tile holds fewer than min_cols columns at full length
                // reads row segments of under 128 B; the four-step form reads whole 16-64
                // column rows in both of its passes, which beats the one short-segment pass.
                uint32_t wfit = g.W;
                while (wfit > 1 && !fits_single(p, n, wfit)) {
                    wfit /= 2;
                }
                can_split = wfit < min_cols(p) && split_two(p, n, min_cols(p), s1, s2);
            }
            if (!can_split && n <= static_cast<uint64_t>(max_cap(p)) && fits_single(p, n, g.W)) {
                p.launches.push_back(make_launch(p, SFH_MI_PLAIN, g, src, dst, false, "single"));
                return;
            }
        }
        // ---------------- two launches (four-step) ----------------
        uint64_t n1 = 0, n2 = 0;
        uint32_t wmin = min_cols(p);
        while (!split_two(p, n, wmin, n1, n2)) {
            if (wmin == 1) {
                throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "transform length too large");
            }
            wmin /= 2;
        }
        const uint64_t chunk = chunk_lines(p, outer, n * inner);
        p.scratch_elems = std::max(p.scratch_elems, chunk * n * inner);
        for (uint64_t o0 = 0; o0 < outer; o0 += chunk) {
            const uint64_t cnt = std::min(chunk, outer - o0);
            const int64_t off = static_cast<int64_t>(o0 * n * inner);
            plan_fourstep(p, cnt, n, inner, n1, n2, src, dst, off);
        }
        return;
    }
    plan_bluestein(p, outer, n, inner, src, dst);
}

// Two launches over `outer` lines starting at element `off` of src/dst (scratch from 0).
static void plan_fourstep(plan &p, uint64_t outer, uint64_t n, uint64_t inner, uint64_t n1, uint64_t n2,
                          buf_id src, buf_id dst, int64_t off)
{
        {
            // pass A: FFT over j1 (n1 points, stride n2*inner) for every (outer, j2, inner),
            // twiddle W_n^(j2 k1), in place positions into scratch.
            geometry g;
            g.L = static_cast<uint32_t>(n1);
            g.wcount = n2 * inner;
            g.W = choose_w(p, n1, g.wcount, true);
            g.nblo = outer;
            g.unit_dim = SFH_UNIT_BLO;
            g.in_w = g.out_w = 1;
            g.in_j = g.out_k = static_cast<int64_t>(n2 * inner);
            g.in_blo = g.out_blo = static_cast<int64_t>(n * inner);
            g.col_div = inner;
            g.col_mod = n2;
            launch l = make_launch(p, SFH_MI_T4, g, src, BUF_SCRATCH, false, "fourstep-A", n);
            l.src_off = off;
            p.launches.push_back(l);
        }
        {
            // pass B: FFT over j2 (n2 points) of row k1, output k1 + n1 k2.
            geometry g;
            g.L = static_cast<uint32_t>(n2);
            if (inner == 1) {
                g.wcount = n1;
                g.W = choose_w(p, n2, n1, true);
                g.nblo = outer;
                g.unit_dim = SFH_UNIT_BLO;
                g.in_w = static_cast<int64_t>(n2);
                g.in_j = 1;
                g.out_w = 1;
                g.out_k = static_cast<int64_t>(n1);
                g.in_blo = g.out_blo = static_cast<int64_t>(n);
            } else {
                g.wcount = inner;
                g.W = choose_w(p, n2, inner, true);
                g.nblo = n1;
                g.nbhi = outer;
                g.unit_dim = SFH_UNIT_BHI;
                g.in_w = g.out_w = 1;
                g.in_j = static_cast<int64_t>(inner);
                g.out_k = static_cast<int64_t>(n1 * inner);
                g.in_blo = static_cast<int64_t>(n2 * inner);
                g.out_blo = static_cast<int64_t>(inner);
                g.in_bhi = g.out_bhi = static_cast<int64_t>(n * inner);
            }
            launch l = make_launch(p, SFH_MI_PLAIN, g, BUF_SCRATCH, dst, false, "fourstep-B");
            l.dst_off = off;
            // Last tiles first: pass A wrote them last, so part of them is still in the
            // 256 MiB Infinity Cache when pass B starts (measured 3-5% at 2^19-2^20 points,
            // neutral below).
            l.args.flags |= SFH_TILE_REVERSE;
            l.desc += " reverse";
            p.launches.push_back(l);
        }
}

// Bluestein with M = 2 Mh >= 2n over two workgroups per transform, for batches too small to
// fill the device with one workgroup per transform.  Since n <= Mh, the zero-padded
// a = x c has an all-zero upper half, so the even and odd halves of DFT_M(a) are the
// length-Mh DFTs of a and of a W_M^j (j < Mh).  Half h runs the convolution for its half of
// the spectrum, X_h[k] = DFT_M(a)[2k + h], in one workgroup:
//   E_h = IDFT_Mh(X_h * S[2k + h])  with S = DFT_M(b) / M,
// and the inverse splits the same way, z[m] = E_0[m] + W_M^-m E_1[m] for m < n.  Launch 1
// writes P_h[m] = c[m] E_0[m] and c[m] W_M^-m E_1[m] (m < n) to scratch; launch 2, a radix-2
// tile of the pair whose difference output is skipped (POST bound 1), writes
// y[m] = P_0[m] + P_1[m].
// Half length of the Bluestein pair for n: the smallest 7-smooth Mh >= n that fits a tile,
// or a longer one (at most 3% longer) that has a length-specialized pair kernel with fewer
// passes; 0 if none fits.
static uint64_t pair_half_length(const plan &p, uint64_t n, const sfh_fixed_kernel **kernel)
{
    const uint64_t cap = static_cast<uint64_t>(max_cap(p));
    const uint64_t Mh = smooth_at_least(n, [&](uint64_t m) { return m <= cap && fits_single(p, m, 1); });
    *kernel = Mh ? find_fixed(p, SFH_FX_BLUE_PAIR, Mh) : nullptr;
    int nk = 0;
    const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
    for (int i = 0; i < nk && Mh; ++i) {
        const sfh_fixed_kernel *k = &ks[i];
        if (k->mode_index == SFH_FX_BLUE_PAIR && k->is_double == (p.dbl ? 1 : 0) && k->L >= n && k->L > Mh
            && 100 * static_cast<uint64_t>(k->L) <= 103 * Mh && k->lds_bytes <= k_lds_limit
            && (!*kernel || k->n1 < (*kernel)->n1) && find_fixed(p, SFH_FX_BLUE_PAIR, k->L) == k) {
            *kernel = k;
            return k->L;
        }
    }
    return Mh;
}

static bool plan_bluestein_halves(plan &p, uint64_t outer, uint64_t n, buf_id src, buf_id dst)
{
    const sfh_fixed_kernel *pair = nullptr;
    const uint64_t Mh = pair_half_length(p, n, &pair);
    if (!Mh) {
        return false;
    }
    const uint64_t M = 2 * Mh;
    std::vector<cxl> chirp(n), bvec(M, cxl{0.0L, 0.0L});
    for (uint64_t j = 0; j < n; ++j) {
        chirp[j] = chirp_ld(j, n);
        const cxl b{chirp[j].re, -chirp[j].im};
        bvec[j] = b;
        if (j) {
            bvec[M - j] = b;
        }
    }
    std::vector<cxl> spec = host_dft_ld(bvec);
    const long double invM = 1.0L / static_cast<long double>(M);
    for (auto &v : spec) {
        v.re *= invM;
        v.im *= invM;
    }
    const auto mul = [](cxl a, cxl b) { return cxl{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; };
    // Interleaved by half so that one bound gives both the zero padding and the truncation:
    // pre[2j + h] = c[j] W_M^(h j) is read iff 2j + h < 2n, i.e. j < n;
    // post[2m + h] = conj(c[m] W_M^(-h m)) (conj-folded) is written iff m < n.
    std::vector<cxl> pre(2 * n), post(2 * n);
    for (uint64_t j = 0; j < n; ++j) {
        pre[2 * j] = chirp[j];
        pre[2 * j + 1] = mul(chirp[j], root_ld(j, M));
        post[2 * j] = cxl{chirp[j].re, -chirp[j].im};
        const cxl c1 = mul(chirp[j], root_ld(M - j, M)); // c[m] W_M^-m
        post[2 * j + 1] = cxl{c1.re, -c1.im};
    }
    const std::string nkey = std::to_string(n) + "_" + std::to_string(M);
    const int64_t pre_off = table(p, "hpre" + nkey, pre);
    const int64_t post_off = table(p, "hpost" + nkey, post);
    const int64_t spec_off = table(p, "spec" + nkey, spec);
    p.scratch_elems = std::max(p.scratch_elems, outer * 2 * n);
    if (pair) {
        // one launch: the second half of each pair to finish adds the two partial results
        geometry g;
        g.L = static_cast<uint32_t>(Mh);
        g.W = 1;
        g.wcount = outer;
        g.unit_dim = SFH_UNIT_COL;
        g.nblo = 2; // the half h
        g.in_w = static_cast<int64_t>(n);
        g.in_j = 1;
        g.in_blo = 0; // both halves read the same transform
        g.out_w = static_cast<int64_t>(n);
        g.out_k = 1;
        g.out_blo = 0; // and write the same output
        g.col_div = 1;
        g.col_mod = 1;
        g.col_bmul = 1; // col = h
        launch l = make_launch(p, SFH_MI_BLUE_FUSED, g, src, dst, true, "bluestein-pair");
        sfh_tile_args &a = l.args;
        a.pre_jmul = 2;
        a.pre_cmul = 1;
        a.nvalid_in = static_cast<uint32_t>(2 * n);
        a.mid_kmul = 2;
        a.mid_cmul = 1;
        a.post_kmul = 2;
        a.post_cmul = 1;
        a.nvalid_out = static_cast<uint32_t>(2 * n);
        a.pair_stride = static_cast<uint32_t>(n);
        l.conj_post = true;
        l.pre_off = pre_off;
        l.mid_off = spec_off;
        l.post_off = post_off;
        l.fixed = pair;
        l.fx_tw_off = fixed_tables(p, *pair);
        l.desc += " fixed=" + std::to_string(pair->nt) + "x" + std::to_string(pair->L) + " pair";
        p.pair_transforms = std::max(p.pair_transforms, outer);
        p.launches.push_back(l);
        return true;
    }
    const int64_t one_off = table(p, "one", std::vector<cxl>{cxl{1.0L, 0.0L}});
    {
        geometry g;
        g.L = static_cast<uint32_t>(Mh);
        g.W = 1;
        g.wcount = outer;
        g.unit_dim = SFH_UNIT_COL;
        g.nblo = 2; // the half h
        g.in_w = static_cast<int64_t>(n);
        g.in_j = 1;
        g.in_blo = 0; // both halves read the same transform
        g.out_w = static_cast<int64_t>(2 * n);
        g.out_k = 1;
        g.out_blo = static_cast<int64_t>(n);
        g.col_div = 1;
        g.col_mod = 1;
        g.col_bmul = 1; // col = h
        launch l = make_launch(p, SFH_MI_BLUE_FUSED, g, src, BUF_SCRATCH, true, "bluestein-half");
        sfh_tile_args &a = l.args;
        a.pre_jmul = 2;
        a.pre_cmul = 1;
        a.nvalid_in = static_cast<uint32_t>(2 * n);
        a.mid_kmul = 2;
        a.mid_cmul = 1;
        a.post_kmul = 2;
        a.post_cmul = 1;
        a.nvalid_out = static_cast<uint32_t>(2 * n);
        l.conj_post = true;
        l.pre_off = pre_off;
        l.mid_off = spec_off;
        l.post_off = post_off;
        p.launches.push_back(l);
    }
    {
        geometry g;
        g.L = 2;
        g.W = choose_w(p, 2, n, true);
        g.wcount = n;
        g.nblo = outer;
        g.unit_dim = SFH_UNIT_BLO;
        g.in_w = 1;
        g.in_j = static_cast<int64_t>(n);
        g.in_blo = static_cast<int64_t>(2 * n);
        g.out_w = 1;
        g.out_k = static_cast<int64_t>(n);
        g.out_blo = static_cast<int64_t>(n);
        launch l = make_launch(p, SFH_MI_BLUE_C, g, BUF_SCRATCH, dst, false, "bluestein-sum");
        l.args.post_kmul = 1; // q = k: only the sum output k = 0 of the pair is written
        l.args.post_cmul = 0;
        l.args.nvalid_out = 1;
        l.post_off = one_off;
        p.launches.push_back(l);
    }
    return true;
}

static void plan_bluestein(plan &p, uint64_t outer, uint64_t n, uint64_t inner, buf_id src, buf_id dst)
{
    // ---------------- Bluestein ----------------
    // chirp c[j] = exp(-i pi j^2/n); b[m] = conj(c[m]) for |m| < n wrapped mod M;
    // X[k] = c[k] * IDFT_M(DFT_M(x c) * DFT_M(b))[k]  (src/sycl/sycl.cpp:268-339, 546-568)
    if (inner == 1 && !p.tune.split) {
        // Two workgroups per transform when one per transform would leave at least half of
        // the CUs idle and the length-specialized pair kernel exists (its second workgroup
        // to finish adds the halves in the same launch).  Measured on n = 10007 (fused
        // 20160-point kernel against the pair of 10080-point halves): batch 64 44 -> 32 us,
        // batch 128 44 -> 36 us, batch 256 50 -> 65 us (two rounds of pairs).  The generic
        // form of the halves needs a second launch for the sum and never pays (batch 128:
        // 63 against 55 us), so it runs only on request.  SYCLFFT_HIP_BLUE_HALVES=1 forces
        // the halves, =0 never takes them.
        int force = -1;
        if (const char *e = std::getenv("SYCLFFT_HIP_BLUE_HALVES")) {
            if (std::strcmp(e, "0") == 0) {
                force = 0;
            } else if (std::strcmp(e, "1") == 0) {
                force = 1;
            } else {
                throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "SYCLFFT_HIP_BLUE_HALVES is not 0 or 1");
            }
        }
        bool take = force == 1;
        if (force < 0) {
            const uint64_t cus = static_cast<uint64_t>(std::max(0, sfh_device_cus()));
            const sfh_fixed_kernel *pair = nullptr;
            const uint64_t Mh = pair_half_length(p, n, &pair);
            take = cus > 0 && 2 * outer <= cus && Mh && pair != nullptr;
        }
        if (take && plan_bluestein_halves(p, outer, n, src, dst)) {
            return;
        }
    }
    const uint64_t need = 2 * n - 1;
    const uint64_t cap = static_cast<uint64_t>(max_cap(p));
    const uint64_t fusedM = smooth_at_least(need, [&](uint64_t m) { return m <= cap && fits_single(p, m, 1); });
    uint64_t M = p.tune.split ? 0 : fusedM;
    bool fused = M != 0;
    uint64_t m1 = 0, m2 = 0;
    if (!fused) {
        uint32_t wmin = min_cols(p);
        for (;;) {
            M = smooth_at_least(need, [&](uint64_t m) {
                uint64_t a, b;
                return split_two(p, m, wmin, a, b);
            });
            if (M) {
                break;
            }
            if (wmin == 1) {
                throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "Bluestein length too large");
            }
            wmin /= 2;
        }
        split_two(p, M, wmin, m1, m2);
        // launch B runs two FFTs of m2 points fused into one program, which needs at least
        // two radix passes; a split that short is only reachable by a forced split of a
        // small M, which then stays fused
        if (radix_plan(m2, 16).size() < 2 && fusedM) {
            M = fusedM;
            fused = true;
        }
    }
    // tables
    std::vector<cxl> chirp(n), bvec(M, cxl{0.0L, 0.0L});
    for (uint64_t j = 0; j < n; ++j) {
        chirp[j] = chirp_ld(j, n);
        const cxl b{chirp[j].re, -chirp[j].im};
        bvec[j] = b;
        if (j) {
            bvec[M - j] = b;
        }
    }
    std::vector<cxl> spec = host_dft_ld(bvec);
    const long double invM = 1.0L / static_cast<long double>(M);
    for (auto &s : spec) {
        s.re *= invM;
        s.im *= invM;
    }
    const std::string nkey = std::to_string(n) + "_" + std::to_string(M);
    const int64_t chirp_off = table(p, "chirp" + nkey, chirp);
    // post table of the conj-before-POST stores: conj(v) c = conj(v conj(c))
    std::vector<cxl> chirpc(chirp);
    for (auto &v : chirpc) {
        v.im = -v.im;
    }
    const int64_t chirpc_off = table(p, "chirpc" + nkey, chirpc);

    if (fused) {
        geometry g;
        g.L = static_cast<uint32_t>(M);
        if (inner == 1) {
            g.W = choose_w(p, M, outer, false);
            while (g.W > 1 && !fits_single(p, M, g.W)) {
                g.W /= 2;
            }
            g.wcount = outer;
            g.unit_dim = SFH_UNIT_COL;
            g.in_w = g.out_w = static_cast<int64_t>(n);
            g.in_j = g.out_k = 1;
        } else {
            g.W = 1;
            while (g.W < 16 && fits_single(p, M, g.W * 2) && g.W * 2 <= inner) {
                g.W *= 2;
            }
            g.wcount = inner;
            g.nblo = outer;
            g.unit_dim = SFH_UNIT_BLO;
            g.in_w = g.out_w = 1;
            g.in_j = g.out_k = static_cast<int64_t>(inner);
            g.in_blo = g.out_blo = static_cast<int64_t>(n * inner);
        }
        launch l = make_launch(p, SFH_MI_BLUE_FUSED, g, src, dst, true, "bluestein-fused");
        sfh_tile_args &a = l.args;
        a.pre_jmul = 1;
        a.pre_cmul = 0;
        a.nvalid_in = static_cast<uint32_t>(n);
        a.mid_kmul = 1;
        a.mid_cmul = 0;
        a.post_kmul = 1;
        a.post_cmul = 0;
        a.nvalid_out = static_cast<uint32_t>(n);
        l.conj_post = true;
        l.pre_off = chirp_off;
        l.post_off = chirpc_off;
        l.mid_off = table(p, "spec" + nkey, spec);
        p.launches.push_back(l);
        return;
    }

    // three launches over a (outer, M, inner) scratch, M = m1 * m2, per group of lines
    std::vector<cxl> spec_t(M);
    for (uint64_t k1 = 0; k1 < m1; ++k1) {
        for (uint64_t k2 = 0; k2 < m2; ++k2) {
            spec_t[k1 * m2 + k2] = spec[k1 + m1 * k2];
        }
    }
    const int64_t spect_off = table(p, "spect" + nkey, spec_t);
    const uint64_t all = outer;
    const uint64_t chunk = chunk_lines(p, all, M * inner);
    p.scratch_elems = std::max(p.scratch_elems, chunk * M * inner);
    for (uint64_t o0 = 0; o0 < all; o0 += chunk) {
    outer = std::min(chunk, all - o0);
    const int64_t off = static_cast<int64_t>(o0 * n * inner);
    {
        // A: a[j] = x[j] c[j] (j < n), column FFT over j1 of j = j1 m2 + j2, twiddle W_M^(j2 k1)
        geometry g;
        g.L = static_cast<uint32_t>(m1);
        g.wcount = m2 * inner;
        g.W = choose_w(p, m1, g.wcount, true);
        g.nblo = outer;
        g.unit_dim = SFH_UNIT_BLO;
        g.in_w = 1;
        g.in_j = static_cast<int64_t>(m2 * inner);
        g.in_blo = static_cast<int64_t>(n * inner);
        g.out_w = 1;
        g.out_k = static_cast<int64_t>(m2 * inner);
        g.out_blo = static_cast<int64_t>(M * inner);
        g.col_div = inner;
        g.col_mod = m2;
        launch l = make_launch(p, SFH_MI_BLUE_A, g, src, BUF_SCRATCH, false, "bluestein-A", M);
        l.args.pre_jmul = static_cast<uint32_t>(m2);
        l.args.pre_cmul = 1;
        l.args.nvalid_in = static_cast<uint32_t>(n);
        l.pre_off = chirp_off;
        l.src_off = off;
        p.launches.push_back(l);
    }
    {
        // B: row k1: FFT over j2, * spectrum (stored transposed), conj, FFT, twiddle W_M^(k1 j2)
        geometry g;
        g.L = static_cast<uint32_t>(m2);
        if (inner == 1) {
            g.wcount = m1;
            g.W = choose_w(p, m2, m1, false);
            g.nblo = outer;
            g.unit_dim = SFH_UNIT_BLO;
            g.in_w = g.out_w = static_cast<int64_t>(m2);
            g.in_j = g.out_k = 1;
            g.in_blo = g.out_blo = static_cast<int64_t>(M);
            g.col_div = 1;
            g.col_mod = m1;
        } else {
            g.wcount = inner;
            g.W = choose_w(p, m2, inner, true);
            g.nblo = m1;
            g.nbhi = outer;
            g.unit_dim = SFH_UNIT_BHI;
            g.in_w = g.out_w = 1;
            g.in_j = g.out_k = static_cast<int64_t>(inner);
            g.in_blo = g.out_blo = static_cast<int64_t>(m2 * inner);
            g.in_bhi = g.out_bhi = static_cast<int64_t>(M * inner);
            g.col_div = 1;
            g.col_mod = 1;
            g.col_bmul = 1;
        }
        launch l = make_launch(p, SFH_MI_BLUE_B, g, BUF_SCRATCH, BUF_SCRATCH, true, "bluestein-B", M);
        l.args.mid_kmul = 1;
        l.args.mid_cmul = static_cast<uint32_t>(m2);
        l.mid_off = spect_off;
        p.launches.push_back(l);
    }
    {
        // C: column FFT over k1 -> conv index j = j2 + m2 j1; conj, * c[j], keep j < n
        geometry g;
        g.L = static_cast<uint32_t>(m1);
        g.wcount = m2 * inner;
        g.W = choose_w(p, m1, g.wcount, true);
        g.nblo = outer;
        g.unit_dim = SFH_UNIT_BLO;
        g.in_w = 1;
        g.in_j = static_cast<int64_t>(m2 * inner);
        g.in_blo = static_cast<int64_t>(M * inner);
        g.out_w = 1;
        g.out_k = static_cast<int64_t>(m2 * inner);
        g.out_blo = static_cast<int64_t>(n * inner);
        g.col_div = inner;
        g.col_mod = m2;
        launch l = make_launch(p, SFH_MI_BLUE_C, g, BUF_SCRATCH, dst, false, "bluestein-C");
        l.args.post_kmul = static_cast<uint32_t>(m2);
        l.args.post_cmul = 1;
        l.args.nvalid_out = static_cast<uint32_t>(n);
        l.conj_post = true;
        l.post_off = chirpc_off;
        l.dst_off = off;
        p.launches.push_back(l);
    }
    }
}

// normalization_scale (src/core/provider_loader.cpp:134-147)
static double norm_scale(uint32_t mode, int32_t dir, uint64_t n)
{
    switch (mode) {
    case SYCLFFT_HIP_NORM_NONE:
        return 1.0;
    case SYCLFFT_HIP_NORM_FORWARD:
        return dir == -1 ? 1.0 / static_cast<double>(n) : 1.0;
    case SYCLFFT_HIP_NORM_BACKWARD:
        return dir == 1 ? 1.0 / static_cast<double>(n) : 1.0;
    case SYCLFFT_HIP_NORM_ORTHOGONAL:
        return 1.0 / std::sqrt(static_cast<double>(n));
    default:
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Invalid normalization");
    }
}

static void validate_config(const syclfft_hip_plan_config_v1 *cfg)
{
    if (!cfg) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plan config must be non-null");
    }
    if (cfg->abi_version != SYCLFFT_DEVICE_PROVIDER_ABI_VERSION) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLUGIN_ERROR, "plan config has an incompatible ABI version");
    }
    if (cfg->scalar != SYCLFFT_HIP_SCALAR_FLOAT && cfg->scalar != SYCLFFT_HIP_SCALAR_DOUBLE) {
        throw sfh_error(SYCLFFT_HIP_STATUS_UNSUPPORTED_PRECISION, "scalar must be float or double");
    }
    if (cfg->direction != -1 && cfg->direction != 1) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Invalid FFT direction");
    }
    // checked_element_count (src/core/provider_loader.cpp:112-132)
    if (cfg->rank < 1 || cfg->rank > 3) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "FFT rank must be between 1 and 3");
    }
    if (cfg->batch_count == 0) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "FFT batch count must be non-zero");
    }
    uint64_t count = cfg->batch_count;
    for (uint32_t i = 0; i < cfg->rank; ++i) {
        const uint64_t len = cfg->lengths[i];
        if (len == 0) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "FFT dimensions must be non-zero");
        }
        if (count > UINT64_MAX / len) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "FFT element count overflows size_t");
        }
        count *= len;
    }
    if (cfg->normalization > SYCLFFT_HIP_NORM_ORTHOGONAL) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Invalid normalization");
    }
    if (cfg->precision > SYCLFFT_HIP_PRECISION_SPLIT_FP16) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Invalid precision mode");
    }
    if (cfg->precision != SYCLFFT_HIP_PRECISION_VALU && cfg->scalar != SYCLFFT_HIP_SCALAR_FLOAT) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_UNSUPPORTED_PRECISION, "split-precision MFMA modes apply to float plans");
    }
    if (cfg->reserved[0] || cfg->reserved[1] || cfg->reserved[2]) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "reserved config fields must be zero");
    }
}

static int check_device(int device, std::string &reason)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        reason = "no HIP device is visible";
        return -1;
    }
    int dev = device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) {
            reason = "hipGetDevice failed";
            return -1;
        }
    }
    if (dev >= count) {
        reason = "device ordinal out of range";
        return -1;
    }
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        reason = "hipGetDeviceProperties failed";
        return -1;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        reason = std::string("device architecture ") + prop.gcnArchName
                 + " is not gfx950 (MI355X); kernels are built for gfx950 only";
        return -1;
    }
    return dev;
}

struct device_guard
{
    int saved{-1};
    explicit device_guard(int dev)
    {
        if (hipGetDevice(&saved) != hipSuccess) {
            saved = -1;
        }
        if (saved != dev) {
            hip_check(hipSetDevice(dev), "hipSetDevice");
        }
    }
    ~device_guard()
    {
        int cur = -1;
        if (saved >= 0 && hipGetDevice(&cur) == hipSuccess && cur != saved) {
            (void)hipSetDevice(saved);
        }
    }
};

static void release(plan *p)
{
    if (!p) {
        return;
    }
    sync_last(*p);
    if (p->last) {
        (void)hipEventDestroy(p->last);
    }
    if (p->half) {
        release(p->half);
    }
    if (p->core) {
        release(p->core);
    }
    if (p->lay_scratch) {
        (void)hipFree(p->lay_scratch);
    }
    if (p->real_tw) {
        (void)hipFree(p->real_tw);
    }
    if (p->real_scratch) {
        (void)hipFree(p->real_scratch);
    }
    if (p->scratch) {
        (void)hipFree(p->scratch);
    }
    if (p->tables) {
        (void)hipFree(p->tables);
    }
    if (p->pair_counts) {
        (void)hipFree(p->pair_counts);
    }
    for (int k = 0; k < plan::k_unit_slots; ++k) {
        if (p->unit_free[k]) {
            (void)hipEventSynchronize(p->unit_free[k]);
            (void)hipEventDestroy(p->unit_free[k]);
        }
        if (p->unit_dev[k]) {
            (void)hipFree(p->unit_dev[k]);
        }
        if (p->unit_host[k]) {
            (void)hipHostFree(p->unit_host[k]);
        }
    }
    delete p;
}

// batch_inner: the batch is the innermost dimension (transform b's element x at x B + b, the
// layout of stride B and distance 1), so every axis is planned as a strided axis whose inner
// extent includes the batch, with tiles of adjacent transforms.
// trail > 1: every transform carries an untransformed innermost dimension of that extent
// (C order [batch][lengths...][trail]); the axes are strided by it.  Multi-D real plans use
// it for the leading axes of the half spectrum [..][n/2 + 1].
// Per-pass twiddle tables of a length-specialized kernel: for every pass q >= 1 of its first
// radix program (then of its second), [t - 1][j] = W_{ns r}^(j t), t < r, j < ns.
static int64_t fixed_tables(plan &p, const sfh_fixed_kernel &k)
{
    std::string key = "fx" + std::to_string(k.L);
    std::vector<cxl> v;
    const auto add = [&](const int *r, int n) {
        uint64_t ns = static_cast<uint64_t>(r[0]);
        key += "_";
        for (int q = 0; q < n; ++q) {
            key += std::to_string(r[q]) + ".";
        }
        for (int q = 1; q < n; ++q) {
            const uint64_t R = static_cast<uint64_t>(r[q]);
            for (uint64_t t = 1; t < R; ++t) {
                for (uint64_t j = 0; j < ns; ++j) {
                    v.push_back(root_ld(j * t, ns * R));
                }
            }
            ns *= R;
        }
    };
    add(k.r1, k.n1);
    if (k.n2) {
        add(k.r2, k.n2);
    }
    return table(p, key, v);
}

// A length-specialized kernel of the plan's precision for `mode_index` and length L (null:
// none within the LDS limit; none at all with SYCLFFT_HIP_FIXED=0).
// Set while a real plan builds its half-length complex plan: entries marked half_only
// qualify only then.
static thread_local bool t_real_half = false;

static const sfh_fixed_kernel *find_fixed(const plan &p, int mode_index, uint64_t L)
{
    if (const char *fe = std::getenv("SYCLFFT_HIP_FIXED")) {
        if (std::strcmp(fe, "0") == 0) {
            return nullptr;
        }
    }
    if (p.cfg.precision != SYCLFFT_HIP_PRECISION_VALU) {
        return nullptr;
    }
    int nk = 0;
    const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
    // SYCLFFT_HIP_FIXED_SKIP=<k> passes over the first k qualifying entries, as in choose_fixed
    int skip = 0;
    if (const char *e = std::getenv("SYCLFFT_HIP_FIXED_SKIP")) {
        skip = std::atoi(e);
    }
    for (int i = 0; i < nk; ++i) {
        if (ks[i].is_double == (p.dbl ? 1 : 0) && ks[i].mode_index == mode_index && ks[i].L == L && ks[i].W == 1
            && ks[i].lds_bytes <= k_lds_limit && ks[i].max_tiles_per_cu == 0 && (!ks[i].half_only || t_real_half)) {
            if (skip-- > 0) {
                continue;
            }
            return &ks[i];
        }
    }
    return nullptr;
}

// A column-tile length-specialized kernel of `mode_index` for columns of n points whose
// column count is a multiple of its columns per tile (null: none; none with
// SYCLFFT_HIP_FIXED=0).
static const sfh_fixed_kernel *find_fixed_cols(const plan &p, int mode_index, uint64_t n, uint64_t wcount)
{
    if (const char *fe = std::getenv("SYCLFFT_HIP_FIXED")) {
        if (std::strcmp(fe, "0") == 0) {
            return nullptr;
        }
    }
    if (p.cfg.precision != SYCLFFT_HIP_PRECISION_VALU) {
        return nullptr;
    }
    int nk = 0;
    const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
    int skip = 0;
    if (const char *e = std::getenv("SYCLFFT_HIP_FIXED_SKIP")) {
        skip = std::atoi(e);
    }
    for (int i = 0; i < nk; ++i) {
        const sfh_fixed_kernel &k = ks[i];
        // (a ragged last tile idles W - wcount mod W columns: taken from 8 tiles per row on)
        const uint64_t kw = static_cast<uint64_t>(k.W);
        if (k.is_double == (p.dbl ? 1 : 0) && k.mode_index == mode_index && k.W > 1 && k.L == n
            && (wcount % kw == 0 || wcount >= 8 * kw) && k.lds_bytes <= k_lds_limit && k.max_tiles_per_cu == 0) {
            if (skip-- > 0) {
                continue;
            }
            return &k;
        }
    }
    return nullptr;
}

// Length-specialized kernels (fft_fixed.hip) for launches that fit their contract: one
// column per tile, contiguous points, no LDS load/store loops, one tile per workgroup,
// plain or Bluestein-fused mode, VALU precision.  SYCLFFT_HIP_FIXED=0 keeps every launch on
// the generic kernel.
static void choose_fixed(plan &p)
{
    if (const char *fe = std::getenv("SYCLFFT_HIP_FIXED")) {
        if (std::strcmp(fe, "0") == 0) {
            return;
        }
        if (std::strcmp(fe, "1") != 0) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "SYCLFFT_HIP_FIXED is not 0 or 1");
        }
    }
    if (p.cfg.precision != SYCLFFT_HIP_PRECISION_VALU) {
        return;
    }
    int nk = 0;
    const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
    const uint64_t cus = static_cast<uint64_t>(std::max(0, sfh_device_cus()));
    // SYCLFFT_HIP_FIXED_SKIP=<k>: pass over the first k qualifying entries of each launch
    // (a test and measurement knob: it reaches the entries listed after the preferred one)
    int skip0 = 0;
    if (const char *e = std::getenv("SYCLFFT_HIP_FIXED_SKIP")) {
        skip0 = std::atoi(e);
    }
    for (auto &l : p.launches) {
        int skip = skip0;
        const sfh_tile_args &a = l.args;
        if (l.fixed || a.W != 1 || a.in_j != 1 || a.out_k != 1 || (a.flags & (SFH_LOAD_LDS | SFH_STORE_LDS)) != 0
            || l.grid < a.ntiles || (l.mode != SFH_MI_PLAIN && l.mode != SFH_MI_BLUE_FUSED)) {
            continue;
        }
        for (int i = 0; i < nk; ++i) {
            const sfh_fixed_kernel &k = ks[i];
            if (k.is_double == (p.dbl ? 1 : 0) && k.mode_index == l.mode && k.L == a.L && k.W == 1
                && k.lds_bytes <= k_lds_limit && (!k.half_only || t_real_half)
                && (k.max_tiles_per_cu == 0
                    || (cus > 0 && a.ntiles <= cus * static_cast<uint64_t>(k.max_tiles_per_cu)))) {
                if (skip-- > 0) {
                    continue;
                }
                l.fixed = &k;
                l.fx_tw_off = fixed_tables(p, k);
                l.desc += " fixed=" + std::to_string(k.nt) + "x" + std::to_string(k.L);
                break;
            }
        }
    }
}

// Back to the generic kernel for a launch whose arguments leave the specialized contract.
static void unfix(launch &l)
{
    if (l.fixed_only) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PLANNING_FAILED, "launch has no generic kernel form");
    }
    if (l.fixed) {
        l.fixed = nullptr;
        const size_t at = l.desc.find(" fixed=");
        if (at != std::string::npos) {
            const size_t end = l.desc.find(' ', at + 1);
            l.desc.erase(at, end == std::string::npos ? std::string::npos : end - at);
        }
    }
}

// One launch of l with arguments a (buffers and tile counts already set).
static hipError_t launch_tile(const plan &p, const launch &l, sfh_tile_args &a, uint32_t grid, hipStream_t stream)
{
    if (l.fixed) {
        const uint32_t kw = static_cast<uint32_t>(l.fixed->W);
        if (a.W != kw || (kw == 1 && (a.in_j != 1 || a.out_k != 1)) || (a.flags & (SFH_LOAD_LDS | SFH_STORE_LDS)) != 0) {
            return hipErrorInvalidConfiguration; // outside the specialized kernel's contract
        }
        a.twtab = l.fx_twtab;
        a.lds_bytes = l.fixed->lds_bytes;
        if (l.fixed->mode_index == SFH_FX_BLUE_PAIR) {
            if (!a.pair_buf || !p.pair_counts) {
                return hipErrorInvalidConfiguration;
            }
            a.pair_count = p.pair_counts;
        }
        return static_cast<hipError_t>(sfh_launch_fixed(l.fixed, &a, stream));
    }
    return static_cast<hipError_t>(sfh_launch_tile(&p.variants[l.variant], l.mode, &a, grid, stream));
}

static plan *build_plan(const syclfft_hip_plan_config_v1 *cfg, int dev, const tuning &tune, bool batch_inner = false,
                        uint64_t trail = 1)
{
    std::unique_ptr<plan, void (*)(plan *)> p(new plan, release);
    p->tune = tune;
    p->cfg = *cfg;
    p->device = dev;
    p->dbl = cfg->scalar == SYCLFFT_HIP_SCALAR_DOUBLE;
    p->esize = p->dbl ? 16 : 8;
    p->variants = sfh_variants(p->dbl ? 1 : 0, &p->nvariants);
    // SYCLFFT_HIP_MAX_TILE caps the LDS tile (complex elements): a tuning and test knob that
    // forces the multi-launch paths at small lengths.
    if (const char *cap = std::getenv("SYCLFFT_HIP_MAX_TILE")) {
        p->cap_limit = std::atoi(cap);
    }
    p->transform = 1;
    for (uint32_t i = 0; i < cfg->rank; ++i) {
        p->transform *= cfg->lengths[i];
    }
    p->total = p->transform * cfg->batch_count * trail;
    p->scale = norm_scale(cfg->normalization, cfg->direction, p->transform);

    // axis loop in the reference order (src/sycl/sycl.cpp:692-767)
    size_t first_begin = 0, first_end = 0, last_begin = 0;
    bool any = false;
    for (uint32_t axis = 0; axis < cfg->rank; ++axis) {
        uint64_t inner = (batch_inner ? cfg->batch_count : 1) * trail;
        for (uint32_t k = axis + 1; k < cfg->rank; ++k) {
            inner *= cfg->lengths[k];
        }
        const uint64_t n = cfg->lengths[axis];
        const uint64_t outer = p->total / (n * inner);
        const buf_id src = p->launches.empty() ? BUF_IN : BUF_OUT;
        const size_t begin = p->launches.size();
        plan_axis(*p, outer, n, inner, src, BUF_OUT);
        if (p->launches.size() > begin) {
            if (!any) {
                first_begin = begin;
                first_end = p->launches.size();
                any = true;
            }
            last_begin = begin;
        }
    }
    if (p->launches.empty()) {
        p->copy_only = 1; // every length is 1: out = scale * in
    } else {
        // backward = conj(FFT(conj x)): conjugate where the plan input is read and where the
        // final output is written; the scale goes on every final write
        for (size_t i = first_begin; i < first_end; ++i) {
            if (p->launches[i].src == BUF_IN && cfg->direction == 1) {
                p->launches[i].args.flags |= SFH_CONJ_IN;
            }
        }
        // store epilogue (re * out_sx, im * out_sy) of every launch: the scale and the output
        // conjugation on final writes, and the conj-before-POST sign
        for (size_t i = 0; i < p->launches.size(); ++i) {
            launch &l = p->launches[i];
            const bool final_write = i >= last_begin && l.dst == BUF_OUT;
            const bool conj_out = final_write && cfg->direction == 1;
            const double sc = final_write ? p->scale : 1.0;
            l.args.out_sx = sc;
            l.args.out_sy = (conj_out != l.conj_post) ? -sc : sc;
        }
    }

    // Split-precision MFMA forms (fp32 plans, precision != VALU): every plain or four-step
    // launch with radix-16 passes on a variant that compiles them (fp32, 16 elements per
    // lane, whole waves) runs its radix-16 passes on MFMA.  Other passes (radix 2/3/4/5/7/8/9)
    // and the Bluestein launches stay on the VALU.
    if (cfg->precision != SYCLFFT_HIP_PRECISION_VALU) {
        const bool bf = cfg->precision == SYCLFFT_HIP_PRECISION_SPLIT_BF16;
        for (auto &l : p->launches) {
            if (l.mode != SFH_MI_PLAIN && l.mode != SFH_MI_T4) {
                continue;
            }
            const sfh_kernel_variant &v = p->variants[l.variant];
            const int sm = (l.mode == SFH_MI_PLAIN ? SFH_MI_PLAIN_BF16 : SFH_MI_T4_BF16) + (bf ? 0 : 2);
            const sfh_tile_args &a = l.args;
            bool radix16 = false;
            for (uint32_t k = 0; k < a.npass; ++k) {
                radix16 = radix16 || a.pass[k].radix == 16;
            }
            if (!v.fn[sm] || v.e != 16 || !radix16 || v.nt % 64 != 0) {
                continue;
            }
            l.mode = sm;
            l.desc += bf ? " split=bf16x3" : " split=fp16x3";
            ++p->split_launches;
        }
    }

    device_guard guard(dev);
    // length-specialized kernels first: the generic-only forms below skip their launches
    choose_fixed(*p);
    // Persistent prefetching form for plain and four-step-A launches with more tiles than
    // resident workgroups: SYCLFFT_HIP_PREFETCH=1 enables it (measured slower than one tile
    // per workgroup at full occupancy on the batched shapes, so off by default).
    {
        const char *pe = std::getenv("SYCLFFT_HIP_PREFETCH");
        const bool allow = pe ? std::atoi(pe) != 0 : p->tune.prefetch != 0;
        const int cus = sfh_device_cus();
        for (auto &l : p->launches) {
            if (!allow || cus <= 0 || (l.args.flags & SFH_LOAD_LDS) || l.fixed) {
                continue;
            }
            const int pf_mode = l.mode == SFH_MI_PLAIN ? SFH_MI_PLAIN_PF : (l.mode == SFH_MI_T4 ? SFH_MI_T4_PF : -1);
            if (pf_mode < 0) {
                continue;
            }
            const int per_cu = sfh_resident_groups(&p->variants[l.variant], pf_mode, l.args.lds_bytes);
            const uint64_t cap = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(cus);
            if (per_cu > 0 && l.args.ntiles > cap) {
                l.mode = pf_mode;
                l.grid = static_cast<uint32_t>(cap);
                l.desc += " persistent grid=" + std::to_string(cap);
            }
        }
    }
    // Global-table twins: a launch whose tile data alone would let more workgroups share a
    // CU than tile + LDS twiddle tables do (by the runtime's occupancy count, which includes
    // the twin's register budget) runs the twin that reads its tables from global memory,
    // when its tiles outnumber the resident slots of the LDS-table form.  Plain (FFT-only)
    // launches only.  SYCLFFT_HIP_TW_GLOBAL=1 takes the twin wherever one exists, =0 never.
    {
        int force = -1;
        if (const char *ge = std::getenv("SYCLFFT_HIP_TW_GLOBAL")) {
            if (std::strcmp(ge, "0") == 0) {
                force = 0;
            } else if (std::strcmp(ge, "1") == 0) {
                force = 1;
            } else {
                throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "SYCLFFT_HIP_TW_GLOBAL is not 0 or 1");
            }
        }
        const int cus = sfh_device_cus();
        for (auto &l : p->launches) {
            if (l.fixed) {
                continue;
            }
            const int twin = force == 0 ? -1 : global_table_twin(*p, l.variant);
            if (twin < 0) {
                continue;
            }
            const uint32_t data_bytes = l.args.lds_tab * static_cast<uint32_t>(p->esize);
            // plain one-tile-per-workgroup launches only, also when forced: the fused modes
            // spill at the twin's register budget, and a persistent grid was sized for the
            // LDS-table variant
            if (l.mode != SFH_MI_PLAIN || l.grid < l.args.ntiles) {
                continue;
            }
            if (force != 1) {
                const int r0 = sfh_resident_groups(&p->variants[l.variant], l.mode, l.args.lds_bytes);
                const int r1 = sfh_resident_groups(&p->variants[twin], l.mode, data_bytes);
                if (cus <= 0 || r0 <= 0 || r1 <= r0
                    || l.args.ntiles <= static_cast<uint64_t>(r0) * static_cast<uint64_t>(cus)) {
                    continue;
                }
                // Measured on f64 2049..2560 points: three -> four workgroups per CU wins
                // while it removes a tail round (batch 1024-2048: 1.03-1.21x) and loses up to
                // 8% over many rounds (batch 4096), where the twin's spill and table reads
                // outweigh the residency; one -> two wins at every batch (1.22-1.42x).
                if (r1 < 2 * r0 && l.args.ntiles > 2 * static_cast<uint64_t>(r1) * static_cast<uint64_t>(cus)) {
                    continue;
                }
            }
            l.variant = twin;
            l.args.lds_bytes = data_bytes;
            l.desc += " tw=global lds=" + std::to_string(data_bytes);
        }
    }
    if (!p->host_tables.empty()) {
        hip_check(hipMalloc(&p->tables, p->host_tables.size()), "hipMalloc(tables)");
        p->tables_bytes = p->host_tables.size();
        hip_check(
            hipMemcpy(p->tables, p->host_tables.data(), p->host_tables.size(), hipMemcpyHostToDevice),
            "hipMemcpy(tables)");
    }
    if (p->scratch_elems) {
        hip_check(hipMalloc(&p->scratch, p->scratch_elems * p->esize), "hipMalloc(scratch)");
    }
    if (p->pair_transforms) {
        hip_check(hipMalloc(reinterpret_cast<void **>(&p->pair_counts), p->pair_transforms * sizeof(uint32_t)),
                  "hipMalloc(pair counters)");
        hip_check(hipMemset(p->pair_counts, 0, p->pair_transforms * sizeof(uint32_t)), "hipMemset(pair counters)");
    }
    // The serialization event only orders device work between streams and gates plan
    // destruction, so it needs no system-scope fence: recording it must not write back
    // the caches after every execution (the caller's own events do that where needed).
    hip_check(hipEventCreateWithFlags(&p->last, hipEventDisableTiming | hipEventDisableSystemFence),
              "hipEventCreate");
    auto ptr = [&](int64_t off) -> const void * {
        return off < 0 ? nullptr : static_cast<const unsigned char *>(p->tables) + off;
    };
    for (auto &l : p->launches) {
        l.args.twtab = ptr(l.tw_off);
        l.args.pre = ptr(l.pre_off);
        l.args.mid = ptr(l.mid_off);
        l.args.post = ptr(l.post_off);
        l.fx_twtab = ptr(l.fx_tw_off);
        if (l.fixed && l.fixed->mode_index != SFH_MI_PLAIN && p->tables) {
            // the length-specialized Bluestein kernels read their tables (pass twiddles of two
            // FFTs, chirps, spectrum) from global memory over eight to ten passes: they touch
            // the block at their start (SYCLFFT_HIP_WARM=0: they do not).  Measured on
            // 10007 x 128 in the round-trip step: 48.8 -> 47.5 us per launch; the plain
            // kernels, whose few table lines the first tiles read anyway, lost 0.4-0.8 us to it
            const char *we = std::getenv("SYCLFFT_HIP_WARM");
            if (!we || std::strcmp(we, "0") != 0) {
                l.args.warm = p->tables;
                l.args.warm_bytes = p->tables_bytes;
            }
        }
    }
    p->host_tables.clear();
    p->host_tables.shrink_to_fit();
    for (const auto &l : p->launches) {
        trace(l.desc);
    }
    return p.release();
}

// ------------------------------------------------------------------------------------
// Autotuner: planning_mode::measure times candidate tunings on the device and caches the
// fastest per (device, shape) in a text file ($SYCLFFT_HIP_TUNE_CACHE, default
// ~/.cache/syclfft_hip/tuning_v1.txt, lines "<key>\t<tuning>").  Plans whose shape is in
// the cache use the cached tuning in every planning mode.  SYCLFFT_HIP_TUNE=<tuning> forces
// one (tests, experiments).
// ------------------------------------------------------------------------------------

static std::string tune_cache_path()
{
    if (const char *e = std::getenv("SYCLFFT_HIP_TUNE_CACHE")) {
        return e;
    }
    const char *home = std::getenv("HOME");
    return std::string(home && home[0] ? home : "/tmp") + "/.cache/syclfft_hip/tuning_v1.txt";
}

static std::mutex &tune_mutex()
{
    static std::mutex m;
    return m;
}

static std::map<std::string, std::string> &tune_cache()
{
    static std::map<std::string, std::string> cache;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        if (FILE *f = std::fopen(tune_cache_path().c_str(), "r")) {
            char line[512];
            while (std::fgets(line, sizeof(line), f)) {
                std::string l(line);
                while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) {
                    l.pop_back();
                }
                const auto tab = l.find('\t');
                if (tab != std::string::npos) {
                    cache[l.substr(0, tab)] = l.substr(tab + 1);
                }
            }
            std::fclose(f);
        }
    }
    return cache;
}

static void tune_store(const std::string &key, const std::string &value)
{
    tune_cache()[key] = value;
    const std::string path = tune_cache_path();
    const auto slash = path.rfind('/');
    if (slash != std::string::npos) {
        // create the directory chain (best effort)
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i) {
            if (i == dir.size() || dir[i] == '/') {
                (void)::mkdir(dir.substr(0, i).c_str(), 0755);
            }
        }
    }
    if (FILE *f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, "%s\t%s\n", key.c_str(), value.c_str());
        std::fclose(f);
    }
}

static std::string tune_key(const syclfft_hip_plan_config_v1 *cfg, int dev)
{
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, dev);
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%.6s:%d|%s|%d|%u:%llux%llux%llu|%llu|%u|%u|%u", prop.gcnArchName,
                  prop.multiProcessorCount, cfg->scalar == SYCLFFT_HIP_SCALAR_DOUBLE ? "f64" : "f32",
                  cfg->direction, cfg->rank, static_cast<unsigned long long>(cfg->lengths[0]),
                  static_cast<unsigned long long>(cfg->rank > 1 ? cfg->lengths[1] : 1),
                  static_cast<unsigned long long>(cfg->rank > 2 ? cfg->lengths[2] : 1),
                  static_cast<unsigned long long>(cfg->batch_count), cfg->in_place, cfg->normalization,
                  cfg->precision);
    return buf;
}

static void execute_plan(plan &p, const void *in, void *out, hipStream_t stream, void *const *deps, size_t ndeps,
                         hipEvent_t done);

// Milliseconds per execution of a trial plan, or a negative value if it cannot be timed.
static double time_plan(plan &p)
{
    const uint64_t bytes = p.total * p.esize;
    void *in = nullptr, *out = nullptr;
    hipStream_t st{};
    hipEvent_t e0{}, e1{};
    double ms = -1.0;
    if (hipMalloc(&in, bytes) != hipSuccess || (!p.cfg.in_place && hipMalloc(&out, bytes) != hipSuccess)) {
        (void)hipGetLastError();
    } else if (hipMemset(in, 0, bytes) == hipSuccess && hipStreamCreate(&st) == hipSuccess
               && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        void *o = p.cfg.in_place ? in : out;
        try {
            for (int i = 0; i < 2; ++i) {
                execute_plan(p, in, o, st, nullptr, 0, nullptr);
            }
            const int reps = 5;
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < reps; ++i) {
                execute_plan(p, in, o, st, nullptr, 0, nullptr);
            }
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) == hipSuccess) {
                float t = 0;
                if (hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
                    ms = t / reps;
                }
            }
        } catch (const std::exception &) {
            ms = -1.0;
        }
        (void)hipStreamSynchronize(st);
    }
    sync_last(p);
    if (e0) {
        (void)hipEventDestroy(e0);
    }
    if (e1) {
        (void)hipEventDestroy(e1);
    }
    if (st) {
        (void)hipStreamDestroy(st);
    }
    if (in) {
        (void)hipFree(in);
    }
    if (out) {
        (void)hipFree(out);
    }
    return ms;
}

static std::string describe_plan(const plan &p);

// The launch sequence of a plan description, for the tuner's repeat test.  A
// length-specialized launch (" fixed=") runs its own kernel whatever generic variant it
// replaced, so that variant's fields (threads, elements, LDS) are left out: candidates that
// differ only there run the same kernels and are not timed again.
static std::string launch_key(const std::string &desc)
{
    std::string out, line;
    std::istringstream in(desc);
    while (std::getline(in, line)) {
        if (line.find(" fixed=") != std::string::npos) {
            std::istringstream words(line);
            std::string w, kept;
            while (words >> w) {
                if (w.rfind("nt=", 0) == 0 || w.rfind("e=", 0) == 0 || w.rfind("lds=", 0) == 0 || w.rfind("lp=", 0) == 0) {
                    continue;
                }
                kept += w + ' ';
            }
            line = kept;
        }
        out += line + '\n';
    }
    return out;
}

// Greedy search: the default, each single change, then the pairs of improving changes.
static tuning autotune(const syclfft_hip_plan_config_v1 *cfg, int dev)
{
    std::map<std::string, double> seen; // launch sequence -> ms
    // split-precision plans: a candidate that runs fewer launches on MFMA than the default
    // plan (e.g. a larger variant without radix-16 passes) is not that precision mode
    int base_split = -1;
    auto trial = [&](const tuning &t) -> double {
        std::unique_ptr<plan, void (*)(plan *)> q(build_plan(cfg, dev, t), release);
        if (base_split < 0) {
            base_split = q->split_launches;
        } else if (q->split_launches < base_split) {
            return -1.0;
        }
        // keyed by the launches alone (not the tuning line), so a tuning that yields the same
        // launches as one already timed is not timed again and cannot win on noise
        std::string d = describe_plan(*q);
        d = launch_key(d.substr(0, d.rfind("tuning ")));
        const auto it = seen.find(d);
        if (it != seen.end()) {
            return it->second;
        }
        const double ms = time_plan(*q);
        seen[d] = ms;
        trace("tune " + t.str() + " " + std::to_string(ms) + " ms");
        return ms;
    };
    tuning best{};
    double best_ms = trial(best);
    std::vector<tuning> singles;
    for (int k = 0; k < 5; ++k) {
        tuning t{};
        if (k == 0) {
            t.split = 1;
        } else if (k == 1) {
            t.wshift = -1;
        } else if (k == 2) {
            t.wshift = 1;
        } else if (k == 3) {
            t.prefetch = 1;
        } else {
            t.grow = 1;
        }
        singles.push_back(t);
    }
    std::vector<tuning> better;
    for (const auto &t : singles) {
        double ms = -1;
        try {
            ms = trial(t);
        } catch (const std::exception &) {
            continue;
        }
        if (ms > 0 && (best_ms <= 0 || ms < best_ms)) {
            better.push_back(t);
        }
        if (ms > 0 && (best_ms <= 0 || ms < best_ms * 0.98)) {
            best = t;
            best_ms = ms;
        }
    }
    for (size_t i = 0; i < better.size(); ++i) {
        for (size_t j = i + 1; j < better.size(); ++j) {
            tuning t = better[i];
            const tuning &u = better[j];
            t.split |= u.split;
            t.prefetch |= u.prefetch;
            t.grow |= u.grow;
            if (u.wshift) {
                t.wshift = u.wshift;
            }
            double ms = -1;
            try {
                ms = trial(t);
            } catch (const std::exception &) {
                continue;
            }
            if (ms > 0 && ms < best_ms * 0.98) {
                best = t;
                best_ms = ms;
            }
        }
    }
    return best;
}

static plan *create_plan(const syclfft_hip_plan_config_v1 *cfg)
{
    validate_config(cfg);
    std::string reason;
    const int dev = check_device(cfg->device, reason);
    if (dev < 0) {
        throw sfh_error(SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE, reason);
    }
    tuning t{};
    if (const char *forced = std::getenv("SYCLFFT_HIP_TUNE")) {
        if (!tuning::parse(forced, t)) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "SYCLFFT_HIP_TUNE is not s<0|1>w<-1|0|1>p<0|1>g<0|1>");
        }
    } else {
        device_guard guard(dev);
        const std::string key = tune_key(cfg, dev);
        std::lock_guard<std::mutex> lock(tune_mutex());
        const auto &cache = tune_cache();
        const auto it = cache.find(key);
        if (it != cache.end() && tuning::parse(it->second, t)) {
            // cached
        } else if (cfg->measure) {
            t = autotune(cfg, dev);
            tune_store(key, t.str());
        }
    }
    return build_plan(cfg, dev, t);
}

// Device-pointer check: the whole [ptr, ptr+bytes) range must lie in one device
// allocation visible to the plan's device (replaces the USM check of validate_pointer,
// src/sycl/sycl.cpp:72-85).  A host pointer would fault the GPU, so it is rejected here.
// Under stream capture (hipGraph) the plan's last-execution event is neither waited for
// nor recorded: it would tie the graph to work outside it, and the captured stream order
// already orders the plan's executions inside the graph.
static bool stream_capturing(hipStream_t stream)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st == hipStreamCaptureStatusActive;
}

// Returns the allocation range [*lo, *hi) when the runtime reports it (0, 0 otherwise).
static void validate_pointer(const plan &p, const void *ptr, uint64_t bytes, const char *name,
                             uintptr_t *lo = nullptr, uintptr_t *hi = nullptr)
{
    if (lo) {
        *lo = *hi = 0;
    }
    if (!ptr) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " must be non-null");
    }
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " is not a HIP device allocation");
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " is not device memory");
    }
    if (attr.type == hipMemoryTypeDevice && attr.device != p.device) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_POINTER,
            std::string(name) + " belongs to another device than the plan's");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void *>(ptr)) == hipSuccess && base) {
        const auto b = reinterpret_cast<uintptr_t>(base);
        const auto q = reinterpret_cast<uintptr_t>(ptr);
        if (q < b || q - b + bytes > size) {
            throw sfh_error(
                SYCLFFT_HIP_STATUS_INVALID_POINTER,
                std::string(name) + " allocation is smaller than the transform");
        }
        if (lo) {
            *lo = b;
            *hi = b + size;
        }
    } else {
        (void)hipGetLastError();
    }
}

// The reference's execute() argument checks (src/sycl/sycl.cpp:621-635) plus device-range checks.
static void check_buffers(const plan &p, const void *in, const void *out)
{
    const uint64_t bytes = p.total * p.esize;
    validate_pointer(p, in, bytes, "input");
    validate_pointer(p, out, bytes, "output");
    if (p.cfg.in_place && in != out) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "In-place plan must be executed with execute(inout)");
    }
    if (!p.cfg.in_place && in == out) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place plan requires distinct allocations");
    }
    if (!p.cfg.in_place) {
        const auto a = reinterpret_cast<uintptr_t>(in), b = reinterpret_cast<uintptr_t>(out);
        if (a < b + bytes && b < a + bytes) {
            throw sfh_error(
                SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place input and output overlap");
        }
    }
}

static void execute_real(
    plan &p, const void *in, void *out, hipStream_t stream, void *const *deps, size_t ndeps, hipEvent_t done);

static void execute_layout(
    plan &p, const void *in, const void *in_im, void *out, void *out_im, hipStream_t stream, void *const *deps,
    size_t ndeps, hipEvent_t done);

// The launches of a complex plan over buffers already checked.  `scratch` is the plan's own
// scratch, or caller-provided device memory of at least scratch_elems elements.
static void execute_checked(
    plan &p, const void *in, void *out, hipStream_t stream, void *const *deps, size_t ndeps, hipEvent_t done,
    void *scratch)
{
    std::lock_guard<std::mutex> lock(p.mu);
    device_guard guard(p.device);
    const bool capturing = stream_capturing(stream);
    for (size_t i = 0; i < ndeps; ++i) {
        if (deps[i]) {
            hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(deps[i]), 0), "hipStreamWaitEvent");
        }
    }
    // the previous execution was enqueued on this same stream: stream order already
    // serializes them
    if (!capturing) {
        wait_last(p, stream);
    }
    if (p.copy_only) {
        hip_check(
            static_cast<hipError_t>(
                sfh_launch_copy_scale(in, out, p.total, p.dbl ? 1 : 0, p.scale, stream)),
            "copy_scale launch");
    }
    void *bufs[3] = {const_cast<void *>(in), out, scratch};
    for (auto &l : p.launches) {
        sfh_tile_args a = l.args;
        a.pair_buf = scratch;
        a.in = static_cast<const unsigned char *>(bufs[l.src]) + l.src_off * static_cast<int64_t>(p.esize);
        a.out = static_cast<unsigned char *>(bufs[l.dst]) + l.dst_off * static_cast<int64_t>(p.esize);
        hip_check(launch_tile(p, l, a, l.grid, stream), "tile kernel launch");
    }
    if (!capturing) {
        hip_check(hipEventRecord(p.last, stream), "hipEventRecord(last)");
        set_last(p, stream);
    }
    if (done) {
        hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
    }
}

static void execute_plan(
    plan &p, const void *in, void *out, hipStream_t stream, void *const *deps, size_t ndeps,
    hipEvent_t done)
{
    if (p.real) {
        execute_real(p, in, out, stream, deps, ndeps, done);
        return;
    }
    if (p.layout) {
        if (p.lay.in_planar || p.lay.out_planar) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                            "planar layout plans execute through syclfft_hip_layout_execute");
        }
        execute_layout(p, in, nullptr, out, nullptr, stream, deps, ndeps, done);
        return;
    }
    check_buffers(p, in, out);
    execute_checked(p, in, out, stream, deps, ndeps, done, p.scratch);
}

// Execution of `count` transforms of the plan's shape given by pointer arrays: transform k
// reads ins[k] and writes outs[k] (count <= the plan's batch count).  Every launch that
// touches a caller buffer gets a unit table (one base pointer per unit of its transform
// dimension) in a device slot of the plan's ring; every launch's transform dimension is cut
// to `count` transforms, so scratch holds the first `count` transforms' intermediates.
// `validate` = false: the caller checked the buffers as whole allocations; `ordered` = false:
// the caller orders executions itself, so a plan without scratch (whose executions share no
// memory) does not wait for its previous execution.
static void execute_units(
    plan &p, uint64_t count, const void *const *ins, void *const *outs, hipStream_t stream, void *const *deps,
    size_t ndeps, hipEvent_t done, bool validate = true, bool ordered = true)
{
    if (p.copy_only || p.launches.empty()) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "pointer-array execution needs a plan with transform launches");
    }
    if (count == 0 || count > p.cfg.batch_count) {
        throw sfh_error(
            SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "transform count must be between 1 and the plan's batch count");
    }
    if (!ins || !outs) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, "pointer arrays must be non-null");
    }
    const uint64_t bytes = p.transform * p.esize;
    // every pointer is checked; pointers inside an allocation already checked (the
    // transforms of one contiguous batch) only need the range test
    uintptr_t lo[2] = {0, 0}, hi[2] = {0, 0};
    const auto check = [&](const void *ptr, int side, const char *name) {
        const auto q = reinterpret_cast<uintptr_t>(ptr);
        if (hi[side] && q >= lo[side] && q + bytes <= hi[side]) {
            return;
        }
        validate_pointer(p, ptr, bytes, name, &lo[side], &hi[side]);
    };
    for (uint64_t k = 0; k < count; ++k) {
        if (validate) {
            check(ins[k], 0, "input");
            check(outs[k], 1, "output");
        }
        if (p.cfg.in_place != (ins[k] == outs[k] ? 1u : 0u)) {
            throw sfh_error(
                SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                p.cfg.in_place ? "In-place plan must be executed with execute(inout)"
                               : "Out-of-place plan requires distinct allocations");
        }
        if (!p.cfg.in_place) {
            const auto a = reinterpret_cast<uintptr_t>(ins[k]), b = reinterpret_cast<uintptr_t>(outs[k]);
            if (a < b + bytes && b < a + bytes) {
                throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place input and output overlap");
            }
        }
    }
    // Transforms laid out like a contiguous batch (ins[k] = ins[0] + k transform sizes, same
    // for outs) need no tables: the launches run on ins[0] / outs[0] with their transform
    // dimension cut to `count`.
    bool affine = true;
    for (uint64_t k = 1; k < count && affine; ++k) {
        affine = static_cast<const unsigned char *>(ins[k]) == static_cast<const unsigned char *>(ins[0]) + k * bytes
                 && static_cast<const unsigned char *>(outs[k]) == static_cast<const unsigned char *>(outs[0]) + k * bytes;
    }
    // table layout: per launch, an input table then an output table where the side is a
    // caller buffer; sized for the plan's batch count
    const uint64_t B = p.cfg.batch_count;
    uint64_t cap_entries = 0;
    for (const auto &l : p.launches) {
        if (l.unit_dim == 0 || l.unit_extent % B != 0 || l.src_off != 0 || l.dst_off != 0) {
            throw sfh_error(
                SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                "plan layout does not support pointer-array execution (chunked or unannotated launch)");
        }
        if (l.unit_dim == SFH_UNIT_COL && (l.args.flags & (SFH_LOAD_LDS | SFH_STORE_LDS))) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "column-unit launch with an LDS load/store loop");
        }
        const uint64_t per = l.unit_extent;
        cap_entries += (l.src != BUF_SCRATCH ? per : 0) + (l.dst != BUF_SCRATCH ? per : 0);
    }
    std::lock_guard<std::mutex> lock(p.mu);
    device_guard guard(p.device);
    const bool capturing = stream_capturing(stream);
    if (capturing && !affine) {
        // a captured copy would read the host staging slot when the graph replays, after
        // later executions may have refilled it
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE,
                        "pointer arrays that need unit tables cannot be captured into a graph");
    }
    if (affine) {
        cap_entries = 0;
    }
    if (!affine && !p.unit_bytes) {
        p.unit_bytes = std::max<uint64_t>(cap_entries, 1) * sizeof(void *);
        for (int k = 0; k < plan::k_unit_slots; ++k) {
            hip_check(hipMalloc(&p.unit_dev[k], p.unit_bytes), "hipMalloc(unit tables)");
            hip_check(hipHostMalloc(&p.unit_host[k], p.unit_bytes, hipHostMallocDefault), "hipHostMalloc(unit tables)");
            hip_check(hipEventCreateWithFlags(&p.unit_free[k], hipEventDisableTiming), "hipEventCreate");
        }
    }
    const int slot = affine ? -1 : p.unit_next;
    void **host = nullptr, **dev = nullptr;
    if (!affine) {
        p.unit_next = (p.unit_next + 1) % plan::k_unit_slots;
        if (p.unit_used[slot]) {
            hip_check(hipEventSynchronize(p.unit_free[slot]), "hipEventSynchronize(unit slot)");
        }
        host = static_cast<void **>(p.unit_host[slot]);
        dev = static_cast<void **>(p.unit_dev[slot]);
    }
    struct side_tables
    {
        const void *const *in;
        void *const *out;
    };
    std::vector<side_tables> tabs(p.launches.size(), side_tables{nullptr, nullptr});
    uint64_t pos = 0;
    const auto fill = [&](const launch &l, buf_id b, int64_t stride) -> void ** {
        const void *const *src = (b == BUF_IN) ? ins : reinterpret_cast<const void *const *>(outs);
        const uint64_t units = l.unit_extent / B;
        void **t = host + pos;
        for (uint64_t k = 0; k < count; ++k) {
            const unsigned char *base = static_cast<const unsigned char *>(src[k]);
            for (uint64_t r = 0; r < units; ++r) {
                t[k * units + r] = const_cast<unsigned char *>(base + static_cast<int64_t>(r) * stride * static_cast<int64_t>(p.esize));
            }
        }
        void **d = dev + pos;
        pos += count * units;
        return d;
    };
    for (size_t i = 0; i < p.launches.size() && !affine; ++i) {
        const launch &l = p.launches[i];
        if (l.src != BUF_SCRATCH) {
            tabs[i].in = fill(l, l.src, l.unit_in_stride);
        }
        if (l.dst != BUF_SCRATCH) {
            tabs[i].out = fill(l, l.dst, l.unit_out_stride);
        }
    }
    for (size_t i = 0; i < ndeps; ++i) {
        if (deps[i]) {
            hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(deps[i]), 0), "hipStreamWaitEvent");
        }
    }
    const bool serialize = (ordered || p.scratch != nullptr) && !capturing;
    if (serialize) {
        wait_last(p, stream);
    }
    if (pos) {
        hip_check(hipMemcpyAsync(dev, host, pos * sizeof(void *), hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync(unit tables)");
    }
    for (size_t i = 0; i < p.launches.size(); ++i) {
        const launch &l = p.launches[i];
        sfh_tile_args a = l.args;
        const uint64_t units = l.unit_extent / B;
        const uint64_t ext = count * units;
        const void *user[2] = {ins[0], outs[0]}; // BUF_IN, BUF_OUT bases of the affine case
        a.in = (l.src == BUF_SCRATCH) ? p.scratch : (affine ? user[l.src] : nullptr);
        a.out = (l.dst == BUF_SCRATCH) ? p.scratch : (affine ? const_cast<void *>(user[l.dst]) : nullptr);
        a.pair_buf = p.scratch;
        a.in_units = tabs[i].in;
        a.out_units = tabs[i].out;
        const uint64_t nbhi = a.ntiles / (a.nwt * a.nblo);
        if (l.unit_dim == SFH_UNIT_COL) {
            a.wcount = ext;
            a.nwt = (ext + a.W - 1) / a.W;
            a.ntiles = a.nwt * a.nblo * nbhi;
        } else if (l.unit_dim == SFH_UNIT_BLO) {
            a.nblo = ext;
            a.ntiles = a.nwt * ext * nbhi;
        } else {
            a.ntiles = a.nwt * a.nblo * ext;
        }
        const bool persistent = l.mode == SFH_MI_PLAIN_PF || l.mode == SFH_MI_T4_PF;
        const uint64_t grid = persistent ? std::min<uint64_t>(l.grid, a.ntiles) : a.ntiles;
        hip_check(launch_tile(p, l, a, static_cast<uint32_t>(std::min<uint64_t>(grid, 0xffffffffu)), stream),
                  "tile kernel launch");
    }
    if (!affine) {
        hip_check(hipEventRecord(p.unit_free[slot], stream), "hipEventRecord(unit slot)");
        p.unit_used[slot] = true;
    }
    if (serialize) {
        hip_check(hipEventRecord(p.last, stream), "hipEventRecord(last)");
        set_last(p, stream);
    }
    if (done) {
        hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
    }
}

// ------------------------------------------------------------------------------------
// Work-queue execution: executions of several plans in one persistent launch per kernel
// variant.  Plans whose execution is one plain-mode launch are grouped by variant; each
// group's launch arguments (one per execution) and work items (one per tile) are uploaded
// from a pinned staging ring and run by the variant's tile_fft_queue entry, whose
// workgroups take items from a device counter.  Other plans execute as usual first, on
// the same stream.
// ------------------------------------------------------------------------------------

struct queue_ctx
{
    static constexpr int k_slots = 32;
    std::mutex mu;
    queue_slot slot[k_slots];
    int next{};
};

// One context per device for the life of the process (executions of any plan may still read
// its slots until they complete).
static queue_ctx &queue_context(int dev)
{
    static std::mutex m;
    static std::map<int, queue_ctx *> ctxs;
    std::lock_guard<std::mutex> l(m);
    queue_ctx *&c = ctxs[dev];
    if (!c) {
        c = new queue_ctx();
    }
    return *c;
}

static bool queueable(const plan &p)
{
    // the queue launch adds a 16-byte item slot behind the largest tile
    // (a plan on a length-specialized kernel runs its own launch: the queue form is the
    // generic program, whose results may differ from the plan's in the last bit)
    return !p.copy_only && p.launches.size() == 1 && p.launches[0].mode == SFH_MI_PLAIN
           && !p.launches[0].fixed && p.variants[p.launches[0].variant].fn[SFH_MI_QUEUE] != nullptr
           && p.launches[0].args.lds_bytes + 32u <= k_lds_limit;
}

// Whether `count` transforms of the plan (count < batch: the first `count` of a contiguous
// batch) can join a queue launch: the launch's transform dimension must be annotated as for
// pointer-array execution, so it can be cut to `count`.
static bool queueable_for(const plan &p, uint64_t count)
{
    if (!queueable(p) || count == 0 || count > p.cfg.batch_count) {
        return false;
    }
    if (count == p.cfg.batch_count) {
        return true;
    }
    const launch &l = p.launches[0];
    return l.unit_dim != 0 && l.unit_extent % p.cfg.batch_count == 0 && l.src_off == 0 && l.dst_off == 0;
}

// Cut a launch's transform dimension to `count` of the plan's B transforms.
static void cut_to_count(const launch &l, sfh_tile_args &a, uint64_t count, uint64_t B)
{
    const uint64_t ext = count * (l.unit_extent / B);
    const uint64_t nbhi = a.ntiles / (a.nwt * a.nblo);
    if (l.unit_dim == SFH_UNIT_COL) {
        a.wcount = ext;
        a.nwt = (ext + a.W - 1) / a.W;
        a.ntiles = a.nwt * a.nblo * nbhi;
    } else if (l.unit_dim == SFH_UNIT_BLO) {
        a.nblo = ext;
        a.ntiles = a.nwt * ext * nbhi;
    } else {
        a.ntiles = a.nwt * a.nblo * ext;
    }
}

// counts (optional): entry k runs the first counts[k] transforms of plans[k] (contiguous from
// ins[k] / outs[k]); validate = false: buffers and placement were checked by the caller, and
// every entry must be queueable; ordered = false: no wait for the plans' previous executions
// on other streams (single-launch plans own no scratch; the serving loop's use)
static void execute_queue(plan *const *plans, const void *const *ins, void *const *outs, size_t count,
                          hipStream_t stream, hipEvent_t done, const uint64_t *counts = nullptr,
                          bool validate = true, bool ordered = true)
{
    if (count == 0) {
        if (done) {
            hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
        }
        return;
    }
    if (!plans || !ins || !outs) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plan and buffer lists must be non-null");
    }
    // allocations already validated in this call: buffers carved from one allocation (the
    // usual case for caching allocators) cost one range query for all of them
    std::vector<std::pair<uintptr_t, uintptr_t>> known;
    const auto check = [&](const plan &p, const void *ptr, const char *name) {
        const uint64_t bytes = p.total * p.esize;
        const auto a = reinterpret_cast<uintptr_t>(ptr);
        for (const auto &r : known) {
            if (ptr && a >= r.first && a < r.second && bytes <= r.second - a) {
                return;
            }
        }
        uintptr_t lo = 0, hi = 0;
        validate_pointer(p, ptr, bytes, name, &lo, &hi);
        if (lo < hi) {
            known.emplace_back(lo, hi);
        }
    };
    for (size_t k = 0; k < count; ++k) {
        if (!plans[k]) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE, "plan handle is null");
        }
        const plan &p = *plans[k];
        if (p.device != plans[0]->device) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plans of one queue execution must share a device");
        }
        if (counts && !queueable_for(p, counts[k])) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "entry cannot join a queue launch");
        }
        if (!validate || (!counts && !queueable(p))) {
            continue; // entries that execute directly are checked by their execute path
        }
        check(p, ins[k], "input");
        check(p, outs[k], "output");
        const uint64_t bytes = p.total * p.esize;
        if (p.cfg.in_place && ins[k] != outs[k]) {
            throw sfh_error(
                SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "In-place plan must be executed with execute(inout)");
        }
        if (!p.cfg.in_place) {
            const auto a = reinterpret_cast<uintptr_t>(ins[k]), b = reinterpret_cast<uintptr_t>(outs[k]);
            if (a == b) {
                throw sfh_error(
                    SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place plan requires distinct allocations");
            }
            if (a < b + bytes && b < a + bytes) {
                throw sfh_error(
                    SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place input and output overlap");
            }
        }
    }
    const int dev = plans[0]->device;
    device_guard guard(dev);
    if (stream_capturing(stream)) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE,
                        "queue executions upload their work list from host staging and cannot be captured");
    }
    // (precision, variant) -> executions
    std::map<std::pair<int, int>, std::vector<size_t>> groups;
    for (size_t k = 0; k < count; ++k) {
        const plan &p = *plans[k];
        if (counts || queueable(p)) {
            groups[{p.dbl ? 1 : 0, p.launches[0].variant}].push_back(k);
        } else {
            execute_plan(*plans[k], ins[k], outs[k], stream, nullptr, 0, nullptr);
        }
    }
    if (groups.empty()) {
        if (done) {
            hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
        }
        return;
    }
    // each queued plan once, locked in address order (no lock-order inversion with other callers)
    std::vector<plan *> locked;
    for (const auto &g : groups) {
        for (size_t k : g.second) {
            locked.push_back(plans[k]);
        }
    }
    std::sort(locked.begin(), locked.end());
    locked.erase(std::unique(locked.begin(), locked.end()), locked.end());
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(locked.size());
    for (plan *p : locked) {
        locks.emplace_back(p->mu);
    }
    for (plan *p : locked) {
        if (ordered) {
            wait_last(*p, stream);
        }
    }
    queue_ctx &qc = queue_context(dev);
    std::lock_guard<std::mutex> ql(qc.mu);
    const int cus = sfh_device_cus();
    for (const auto &g : groups) {
        const std::vector<size_t> &ks = g.second;
        const plan &p0 = *plans[ks[0]];
        const sfh_kernel_variant &v = p0.variants[g.first.second];
        // each entry's launch arguments: its buffers, its transform count
        std::vector<sfh_tile_args> eargs(ks.size());
        uint64_t nitems = 0;
        uint32_t lds = 0;
        for (size_t i = 0; i < ks.size(); ++i) {
            const plan &p = *plans[ks[i]];
            const launch &l = p.launches[0];
            void *bufs[3] = {const_cast<void *>(ins[ks[i]]), outs[ks[i]], p.scratch};
            sfh_tile_args &a = eargs[i];
            a = l.args;
            a.in = static_cast<const unsigned char *>(bufs[l.src]) + l.src_off * static_cast<int64_t>(p.esize);
            a.out = static_cast<unsigned char *>(bufs[l.dst]) + l.dst_off * static_cast<int64_t>(p.esize);
            if (counts && counts[ks[i]] < p.cfg.batch_count) {
                cut_to_count(l, a, counts[ks[i]], p.cfg.batch_count);
            }
            nitems += a.ntiles;
            lds = std::max(lds, a.lds_bytes);
        }
        if (nitems > 0xffffffffull) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "queue execution has too many tiles");
        }
        const uint32_t lds_next = (lds + 15u) & ~15u;
        const uint32_t lds_total = lds_next + 16u;
        const size_t args_off = 16;
        const size_t items_off = args_off + ks.size() * sizeof(sfh_tile_args);
        const size_t bytes = items_off + nitems * sizeof(sfh_queue_item);
        // the next slot whose launch has completed; the host waits only when every slot is
        // still in flight
        int pick = -1;
        for (int j = 0; j < queue_ctx::k_slots && pick < 0; ++j) {
            const int c = (qc.next + j) % queue_ctx::k_slots;
            if (!qc.slot[c].used) {
                pick = c;
            } else {
                const hipError_t q = hipEventQuery(qc.slot[c].free);
                if (q == hipSuccess) {
                    pick = c;
                } else if (q != hipErrorNotReady) {
                    hip_check(q, "hipEventQuery(queue slot)");
                }
            }
        }
        if (pick < 0) {
            pick = qc.next % queue_ctx::k_slots;
        }
        qc.next = pick + 1;
        queue_slot &sl = qc.slot[pick];
        if (sl.used) {
            hip_check(hipEventSynchronize(sl.free), "hipEventSynchronize(queue slot)");
        }
        sl.gen.fetch_add(1); // plans of the slot's previous launch now read it as complete
        if (sl.bytes < bytes) {
            if (sl.dev) {
                (void)hipFree(sl.dev);
                (void)hipHostFree(sl.host);
                sl.dev = sl.host = nullptr;
                sl.bytes = 0;
            }
            const size_t want = std::max<size_t>(bytes + bytes / 2, 64 << 10);
            hip_check(hipMalloc(&sl.dev, want), "hipMalloc(queue slot)");
            hip_check(hipHostMalloc(&sl.host, want, hipHostMallocDefault), "hipHostMalloc(queue slot)");
            sl.bytes = want;
        }
        if (!sl.free) {
            hip_check(hipEventCreateWithFlags(&sl.free, hipEventDisableTiming | hipEventDisableSystemFence),
                      "hipEventCreate(queue slot)");
        }
        unsigned char *h = static_cast<unsigned char *>(sl.host);
        std::memset(h, 0, args_off); // the item counter starts at 0
        sfh_tile_args *args = reinterpret_cast<sfh_tile_args *>(h + args_off);
        sfh_queue_item *items = reinterpret_cast<sfh_queue_item *>(h + items_off);
        uint64_t n = 0;
        for (size_t i = 0; i < ks.size(); ++i) {
            std::memcpy(&args[i], &eargs[i], sizeof(sfh_tile_args));
            for (uint64_t t = 0; t < eargs[i].ntiles; ++t) {
                items[n++] = sfh_queue_item{static_cast<uint32_t>(i), static_cast<uint32_t>(t)};
            }
        }
        hip_check(hipMemcpyAsync(sl.dev, sl.host, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(queue)");
        unsigned char *d = static_cast<unsigned char *>(sl.dev);
        sfh_queue_args q{};
        q.args = reinterpret_cast<const sfh_tile_args *>(d + args_off);
        q.items = reinterpret_cast<const sfh_queue_item *>(d + items_off);
        q.counter = reinterpret_cast<uint32_t *>(d);
        q.nitems = static_cast<uint32_t>(nitems);
        q.lds_next = lds_next;
        const int per_cu = std::max(1, sfh_resident_groups(&v, SFH_MI_QUEUE, lds_total));
        const uint64_t grid = std::min<uint64_t>(nitems, static_cast<uint64_t>(per_cu) * std::max(1, cus));
        hip_check(static_cast<hipError_t>(sfh_launch_queue(&v, &q, static_cast<uint32_t>(grid), lds_total, stream)),
                  "queue kernel launch");
        trace("queue launch " + std::string(g.first.first ? "f64" : "f32") + " nt=" + std::to_string(v.nt)
              + " e=" + std::to_string(v.e) + " executions=" + std::to_string(ks.size())
              + " items=" + std::to_string(nitems) + " grid=" + std::to_string(grid)
              + " lds=" + std::to_string(lds_total));
        hip_check(hipEventRecord(sl.free, stream), "hipEventRecord(queue slot)");
        sl.used = true;
        // the slot's event now marks these plans' last execution (no event record per plan)
        const uint64_t gen = sl.gen.load();
        for (size_t k : ks) {
            plan &p = *plans[k];
            p.have_last = true;
            p.last_stream = stream;
            p.last_queue = &sl;
            p.last_qgen = gen;
        }
    }
    if (done) {
        hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
    }
}

// Real 1-D transform: R2C = complex plan of length M over the real input viewed as M complex
// values, then the untangle pass into the M + 1 outputs; C2R = the tangle pass into scratch,
// then the complex backward plan writing the real output viewed as M complex values.
static void execute_real(
    plan &p, const void *in, void *out, hipStream_t stream, void *const *deps, size_t ndeps, hipEvent_t done)
{
    const uint64_t rows = p.transform / (2 * p.real_m); // 1 for 1-D plans
    const uint64_t B = p.cfg.batch_count * rows, M = p.real_m;
    const uint64_t real_bytes = B * M * p.esize;       // B N real values
    const uint64_t half_bytes = B * (M + 1) * p.esize; // B (M + 1) complex values
    const uint64_t in_bytes = p.real == 1 ? real_bytes : half_bytes;
    const uint64_t out_bytes = p.real == 1 ? half_bytes : real_bytes;
    validate_pointer(p, in, in_bytes, "input");
    validate_pointer(p, out, out_bytes, "output");
    const auto a = reinterpret_cast<uintptr_t>(in), b = reinterpret_cast<uintptr_t>(out);
    if (a == b) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place plan requires distinct allocations");
    }
    if (a < b + out_bytes && b < a + in_bytes) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place input and output overlap");
    }
    std::lock_guard<std::mutex> lock(p.mu);
    device_guard guard(p.device);
    const bool capturing = stream_capturing(stream);
    for (size_t i = 0; i < ndeps; ++i) {
        if (deps[i]) {
            hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(deps[i]), 0), "hipStreamWaitEvent");
        }
    }
    if (!capturing) {
        wait_last(p, stream);
    }
    if (p.core) { // multi-D: the last axis as 1-D real rows plus the leading axes
        if (p.real == 1) {
            execute_plan(*p.half, in, out, stream, nullptr, 0, nullptr);
            execute_checked(*p.core, out, out, stream, nullptr, 0, nullptr, p.core->scratch);
        } else {
            execute_checked(*p.core, in, p.real_scratch, stream, nullptr, 0, nullptr, p.core->scratch);
            execute_plan(*p.half, p.real_scratch, out, stream, nullptr, 0, nullptr);
        }
    } else if (p.real_fused) {
        launch &l = p.launches[0];
        sfh_tile_args a = l.args;
        a.in = in;
        a.out = out;
        hip_check(launch_tile(*p.half, l, a, l.grid, stream), p.real == 1 ? "fused r2c launch" : "fused c2r launch");
    } else if (p.real == 1) {
        execute_plan(*p.half, in, p.real_scratch, stream, nullptr, 0, nullptr);
        hip_check(static_cast<hipError_t>(sfh_launch_real_pass(0, p.real_scratch, out, p.real_tw, M, B, p.dbl ? 1 : 0,
                                                               p.scale, stream)),
                  "untangle launch");
    } else {
        hip_check(static_cast<hipError_t>(sfh_launch_real_pass(1, in, p.real_scratch, p.real_tw, M, B, p.dbl ? 1 : 0,
                                                               p.scale, stream)),
                  "tangle launch");
        execute_plan(*p.half, p.real_scratch, out, stream, nullptr, 0, nullptr);
    }
    if (!capturing) {
        hip_check(hipEventRecord(p.last, stream), "hipEventRecord(last)");
        set_last(p, stream);
    }
    if (done) {
        hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
    }
}

static plan *create_real_plan_nd(const syclfft_hip_plan_config_v1 *cfg);

static plan *create_real_plan(const syclfft_hip_plan_config_v1 *cfg)
{
    validate_config(cfg);
    if (cfg->lengths[cfg->rank - 1] < 2 || cfg->lengths[cfg->rank - 1] % 2 != 0) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "real plans need an even last length >= 2");
    }
    if (cfg->in_place) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                        "real plans are out of place (N real values and N/2 + 1 complex values differ in size)");
    }
    if (cfg->precision != SYCLFFT_HIP_PRECISION_VALU) {
        throw sfh_error(SYCLFFT_HIP_STATUS_UNSUPPORTED_PRECISION, "real plans use the VALU precision mode");
    }
    if (cfg->rank > 1) {
        return create_real_plan_nd(cfg);
    }
    syclfft_hip_plan_config_v1 h = *cfg;
    h.lengths[0] = cfg->lengths[0] / 2;
    h.normalization = SYCLFFT_HIP_NORM_NONE;
    struct half_scope
    {
        half_scope() { t_real_half = true; }
        ~half_scope() { t_real_half = false; }
    };
    std::unique_ptr<plan, void (*)(plan *)> hp(nullptr, release);
    {
        const half_scope scope;
        hp.reset(create_plan(&h));
    }
    std::unique_ptr<plan, void (*)(plan *)> p(new plan, release);
    p->cfg = *cfg;
    p->device = hp->device;
    p->dbl = hp->dbl;
    p->esize = hp->esize;
    p->variants = hp->variants;
    p->nvariants = hp->nvariants;
    p->tune = hp->tune;
    p->transform = cfg->lengths[0];
    p->total = p->transform * cfg->batch_count;
    p->scale = norm_scale(cfg->normalization, cfg->direction, p->transform);
    p->real = cfg->direction == -1 ? 1 : 2;
    p->real_m = h.lengths[0];
    const uint64_t M = p->real_m, N = cfg->lengths[0];
    device_guard guard(p->device);
    std::vector<unsigned char> host((M + 1) * p->esize);
    for (uint64_t k = 0; k <= M; ++k) {
        const cxl w = root_ld(k, N);
        if (p->dbl) {
            double *d = reinterpret_cast<double *>(host.data()) + 2 * k;
            d[0] = static_cast<double>(w.re);
            d[1] = static_cast<double>(w.im);
        } else {
            float *f = reinterpret_cast<float *>(host.data()) + 2 * k;
            f[0] = static_cast<float>(w.re);
            f[1] = static_cast<float>(w.im);
        }
    }
    hip_check(hipMalloc(&p->real_tw, host.size()), "hipMalloc(real twiddles)");
    hip_check(hipMemcpy(p->real_tw, host.data(), host.size(), hipMemcpyHostToDevice), "hipMemcpy(real twiddles)");
    hip_check(hipMalloc(&p->real_scratch, cfg->batch_count * M * p->esize), "hipMalloc(real scratch)");
    hip_check(hipEventCreateWithFlags(&p->last, hipEventDisableTiming | hipEventDisableSystemFence), "hipEventCreate");
    // C2R: when the half-length plan is one length-specialized launch and a C2R form of the
    // same radix program exists, the tangle moves into that launch's loads (SYCLFFT_HIP_FIXED=0
    // keeps the separate pass, as it keeps the generic kernels)
    if (p->real == 2 && hp->launches.size() == 1 && !hp->copy_only && hp->launches[0].fixed
        && hp->launches[0].fixed->W == 1 && hp->launches[0].fixed->mode_index == SFH_MI_PLAIN
        && hp->launches[0].args.in_w == static_cast<int32_t>(M)) {
        const sfh_fixed_kernel &pk = *hp->launches[0].fixed;
        int nk = 0;
        const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
        for (int i = 0; i < nk; ++i) {
            const sfh_fixed_kernel &k = ks[i];
            if (k.mode_index != SFH_FX_C2R || k.is_double != pk.is_double || k.L != pk.L || k.nt != pk.nt
                || k.n1 != pk.n1 || !std::equal(pk.r1, pk.r1 + pk.n1, k.r1)) {
                continue;
            }
            launch l = hp->launches[0]; // pass tables, epilogue (conjugation) of the backward plan
            l.fixed = &k;
            l.args.pre = p->real_tw;
            l.args.in_w = static_cast<int32_t>(M + 1);
            l.args.lds_bytes = k.lds_bytes;
            l.args.out_sx *= p->scale;
            l.args.out_sy *= p->scale;
            l.desc = "c2r-fused m=" + std::to_string(M) + " " + l.desc;
            p->launches.push_back(l);
            p->real_fused = true;
            break;
        }
    }
    // R2C: the same with the untangle in that launch's stores (L + 1 outputs per transform)
    if (p->real == 1 && hp->launches.size() == 1 && !hp->copy_only && hp->launches[0].fixed
        && hp->launches[0].fixed->W == 1 && hp->launches[0].fixed->mode_index == SFH_MI_PLAIN
        && hp->launches[0].args.in_w == static_cast<int32_t>(M)
        && hp->launches[0].args.out_w == static_cast<int32_t>(M)) {
        const sfh_fixed_kernel &pk = *hp->launches[0].fixed;
        int nk = 0;
        const sfh_fixed_kernel *ks = sfh_fixed_kernels(&nk);
        for (int i = 0; i < nk; ++i) {
            const sfh_fixed_kernel &k = ks[i];
            if (k.mode_index != SFH_FX_R2C || k.is_double != pk.is_double || k.L != pk.L || k.nt != pk.nt
                || k.n1 != pk.n1 || !std::equal(pk.r1, pk.r1 + pk.n1, k.r1)) {
                continue;
            }
            launch l = hp->launches[0];
            l.fixed = &k;
            l.args.pre = p->real_tw;
            l.args.out_w = static_cast<int32_t>(M + 1);
            l.args.lds_bytes = k.lds_bytes;
            l.args.out_sx *= p->scale;
            l.args.out_sy *= p->scale;
            l.desc = "r2c-fused m=" + std::to_string(M) + " " + l.desc;
            p->launches.push_back(l);
            p->real_fused = true;
            break;
        }
    }
    p->half = hp.release();
    trace(std::string(p->real == 1 ? "real r2c" : "real c2r") + " n=" + std::to_string(N) + " over complex n=" +
          std::to_string(M));
    return p.release();
}


// Elements of one side of a layout: the highest offset touched plus one.
static uint64_t layout_extent(uint64_t batch, uint64_t transform, uint64_t stride, uint64_t dist)
{
    return (batch - 1) * dist + (transform - 1) * stride + 1;
}

// Layout plan: the contiguous plan of the config (`core`) plus the caller's strides and
// planar flags.  A rank-1 interleaved layout whose core is one tile launch reading the input
// and writing the output as contiguous columns (element j of transform b at b n + j) gets
// the layout's distance and stride written into that launch's column and element strides,
// so the launch reads and writes the caller's buffers directly; its tile, lane mapping and
// LDS layout stay those of the contiguous plan.  Every other layout gathers into scratch,
// runs the core in place there and scatters.
static plan *create_layout_plan(const syclfft_hip_plan_config_v1 *cfg, const syclfft_hip_layout_v1 *lay)
{
    validate_config(cfg);
    if (!lay) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "layout is null");
    }
    const uint64_t lim = uint64_t(1) << 40;
    for (uint64_t v : {lay->in_stride, lay->in_distance, lay->out_stride, lay->out_distance}) {
        if (v == 0 || v >= lim) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "layout strides and distances must be in [1, 2^40)");
        }
    }
    if (lay->in_planar > 1 || lay->out_planar > 1) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "layout planar flags must be 0 or 1");
    }
    uint64_t T = 1;
    for (uint32_t d = 0; d < cfg->rank; ++d) {
        T *= cfg->lengths[d];
    }
    const uint64_t B = cfg->batch_count;
    if (layout_extent(B, T, lay->in_stride, lay->in_distance) >= (uint64_t(1) << 62)
        || layout_extent(B, T, lay->out_stride, lay->out_distance) >= (uint64_t(1) << 62)) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "layout extent exceeds 2^62 elements");
    }
    const bool same = lay->in_stride == lay->out_stride && lay->in_distance == lay->out_distance
                      && lay->in_planar == lay->out_planar;
    const bool interleaved = !lay->in_planar && !lay->out_planar;
    std::unique_ptr<plan, void (*)(plan *)> core(nullptr, release);
    bool fused = false;
    if (interleaved && same && B > 1 && lay->in_stride == B && lay->in_distance == 1 && T > 1) {
        // interleaved batches: the contiguous plan of the transposed problem, whose axes are
        // strided by B with adjacent transforms in each tile, reads and writes the caller's
        // buffers directly (every plan form, any rank)
        std::string reason;
        const int dev = check_device(cfg->device, reason);
        if (dev < 0) {
            throw sfh_error(SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE, reason);
        }
        tuning t{};
        if (const char *forced = std::getenv("SYCLFFT_HIP_TUNE")) {
            (void)tuning::parse(forced, t);
        }
        core.reset(build_plan(cfg, dev, t, true));
        for (auto &l : core->launches) {
            l.desc += " layout(batch-inner " + std::to_string(B) + ")";
        }
        fused = true;
    }
    if (!fused && cfg->rank == 1 && interleaved && (!cfg->in_place || same) && T > 1) {
        core.reset(create_plan(cfg));
        plan &c = *core;
        const uint64_t n = T;
        if (!c.copy_only && c.launches.size() == 1) {
            launch &l = c.launches[0];
            sfh_tile_args &a = l.args;
            // a launch only a length-specialized kernel can run keeps contiguous points: strided
            // layouts go through the gather / scatter path
            const bool contiguous = l.src == BUF_IN && l.dst == BUF_OUT && l.src_off == 0 && l.dst_off == 0
                                    && a.nblo == 1 && a.in_j == 1 && a.out_k == 1
                                    && (!l.fixed_only || (lay->in_stride == 1 && lay->out_stride == 1))
                                    && static_cast<uint64_t>(a.in_w) == n && static_cast<uint64_t>(a.out_w) == n
                                    && a.in_units == nullptr && a.out_units == nullptr;
            const auto span = [&](uint64_t sw, uint64_t sj) { return sw * (a.W - 1) + sj * (a.L - 1); };
            const uint64_t i31 = uint64_t(1) << 31;
            if (contiguous && lay->in_distance < i31 && lay->out_distance < i31 && lay->in_stride < i31
                && lay->out_stride < i31 && span(lay->in_distance, lay->in_stride) < i31
                && span(lay->out_distance, lay->out_stride) < i31) {
                a.in_w = static_cast<int32_t>(lay->in_distance);
                a.in_j = static_cast<int32_t>(lay->in_stride);
                a.out_w = static_cast<int32_t>(lay->out_distance);
                a.out_k = static_cast<int32_t>(lay->out_stride);
                if (a.in_j != 1 || a.out_k != 1) {
                    unfix(l); // strided points: the generic kernel
                }
                l.desc += " layout(in " + std::to_string(lay->in_stride) + "/" + std::to_string(lay->in_distance)
                          + ", out " + std::to_string(lay->out_stride) + "/" + std::to_string(lay->out_distance) + ")";
                fused = true;
            }
        }
        if (!fused) {
            core.reset();
        }
    }
    if (!fused) {
        syclfft_hip_plan_config_v1 cc = *cfg;
        cc.in_place = 1;
        core.reset(create_plan(&cc));
    }
    std::unique_ptr<plan, void (*)(plan *)> p(new plan, release);
    p->cfg = *cfg;
    p->device = core->device;
    p->dbl = core->dbl;
    p->esize = core->esize;
    p->variants = core->variants;
    p->nvariants = core->nvariants;
    p->tune = core->tune;
    p->transform = T;
    p->total = T * B;
    p->scale = core->scale;
    p->layout = 1;
    p->lay = *lay;
    p->lay_fused = fused;
    device_guard guard(p->device);
    if (!fused) {
        hip_check(hipMalloc(&p->lay_scratch, p->total * p->esize), "hipMalloc(layout scratch)");
    }
    hip_check(hipEventCreateWithFlags(&p->last, hipEventDisableTiming | hipEventDisableSystemFence), "hipEventCreate");
    p->core = core.release();
    trace(std::string("layout plan ") + (fused ? "fused" : "gather-scatter") + " in " + std::to_string(lay->in_stride)
          + "/" + std::to_string(lay->in_distance) + (lay->in_planar ? " planar" : "") + " out "
          + std::to_string(lay->out_stride) + "/" + std::to_string(lay->out_distance)
          + (lay->out_planar ? " planar" : ""));
    return p.release();
}

static void execute_layout(
    plan &p, const void *in, const void *in_im, void *out, void *out_im, hipStream_t stream, void *const *deps,
    size_t ndeps, hipEvent_t done)
{
    const syclfft_hip_layout_v1 &L = p.lay;
    const uint64_t B = p.cfg.batch_count, T = p.transform;
    if (static_cast<bool>(in_im) != static_cast<bool>(L.in_planar)
        || static_cast<bool>(out_im) != static_cast<bool>(L.out_planar)) {
        throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_POINTER,
                        "imaginary-part pointers are required for planar sides and must be null otherwise");
    }
    struct range
    {
        uintptr_t a, b;
    };
    std::vector<range> ins, outs;
    const auto side = [&](const void *re, const void *im, uint64_t stride, uint64_t dist, uint32_t planar,
                          const char *name, const char *name_im, std::vector<range> &rs) {
        const uint64_t bytes = layout_extent(B, T, stride, dist) * (planar ? p.esize / 2 : p.esize);
        validate_pointer(p, re, bytes, name);
        rs.push_back({reinterpret_cast<uintptr_t>(re), reinterpret_cast<uintptr_t>(re) + bytes});
        if (planar) {
            validate_pointer(p, im, bytes, name_im);
            rs.push_back({reinterpret_cast<uintptr_t>(im), reinterpret_cast<uintptr_t>(im) + bytes});
            if (rs[0].a < rs[1].b && rs[1].a < rs[0].b) {
                throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                                std::string(name) + " real and imaginary arrays overlap");
            }
        }
    };
    side(in, in_im, L.in_stride, L.in_distance, L.in_planar, "input", "input imaginary part", ins);
    side(out, out_im, L.out_stride, L.out_distance, L.out_planar, "output", "output imaginary part", outs);
    if (p.cfg.in_place) {
        if (in != out || in_im != out_im) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "In-place plan must be executed with execute(inout)");
        }
    } else {
        for (const range &x : ins) {
            for (const range &y : outs) {
                if (x.a < y.b && y.a < x.b) {
                    throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "Out-of-place input and output overlap");
                }
            }
        }
    }
    std::lock_guard<std::mutex> lock(p.mu);
    device_guard guard(p.device);
    const bool capturing = stream_capturing(stream);
    for (size_t i = 0; i < ndeps; ++i) {
        if (deps[i]) {
            hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(deps[i]), 0), "hipStreamWaitEvent");
        }
    }
    if (!capturing) {
        wait_last(p, stream);
    }
    const int dbl = p.dbl ? 1 : 0;
    if (p.lay_fused) {
        execute_checked(*p.core, in, out, stream, nullptr, 0, nullptr, p.core->scratch);
    } else {
        const int64_t t = static_cast<int64_t>(T);
        hip_check(static_cast<hipError_t>(sfh_launch_layout_copy(
                      in, in_im, static_cast<int64_t>(L.in_stride), static_cast<int64_t>(L.in_distance),
                      static_cast<int>(L.in_planar), p.lay_scratch, nullptr, 1, t, 0, T, B, dbl, stream)),
                  "layout gather launch");
        execute_checked(*p.core, p.lay_scratch, p.lay_scratch, stream, nullptr, 0, nullptr, p.core->scratch);
        hip_check(static_cast<hipError_t>(sfh_launch_layout_copy(
                      p.lay_scratch, nullptr, 1, t, 0, out, out_im, static_cast<int64_t>(L.out_stride),
                      static_cast<int64_t>(L.out_distance), static_cast<int>(L.out_planar), T, B, dbl, stream)),
                  "layout scatter launch");
    }
    if (!capturing) {
        hip_check(hipEventRecord(p.last, stream), "hipEventRecord(last)");
        set_last(p, stream);
    }
    if (done) {
        hip_check(hipEventRecord(done, stream), "hipEventRecord(done)");
    }
}


// Multi-D real plan, lengths n_0 .. n_{r-1} (n_{r-1} = N even), half spectrum h = N/2 + 1 on
// the last axis, C order [batch][n_0]..[n_{r-2}][h]:
//   R2C: the 1-D real plan of length N over batch * n_0 ... n_{r-2} rows, writing the half
//        spectrum; then the complex plan over axes 0 .. r-2 in place on it, with h as an
//        untransformed innermost extent (every axis strided by h, tiles of adjacent bins);
//   C2R: that complex plan (backward) from the input into scratch, then the 1-D C2R plan.
// Each part normalizes by its own lengths with the plan's mode, which multiply to the
// normalization by prod(n).
static plan *create_real_plan_nd(const syclfft_hip_plan_config_v1 *cfg)
{
    const uint32_t r = cfg->rank;
    const uint64_t N = cfg->lengths[r - 1], hb = N / 2 + 1;
    uint64_t rows = 1;
    for (uint32_t d = 0; d + 1 < r; ++d) {
        rows *= cfg->lengths[d];
    }
    syclfft_hip_plan_config_v1 c1 = *cfg; // the last axis, as 1-D real rows
    c1.rank = 1;
    c1.lengths[0] = N;
    c1.lengths[1] = c1.lengths[2] = 0;
    c1.batch_count = cfg->batch_count * rows;
    std::unique_ptr<plan, void (*)(plan *)> last(create_real_plan(&c1), release);
    syclfft_hip_plan_config_v1 cr = *cfg; // the leading axes of the half spectrum
    cr.rank = r - 1;
    for (uint32_t d = 0; d < 3; ++d) {
        cr.lengths[d] = d + 1 < r ? cfg->lengths[d] : 0;
    }
    cr.in_place = cfg->direction == -1 ? 1u : 0u;
    cr.measure = 0;
    tuning t{};
    if (const char *forced = std::getenv("SYCLFFT_HIP_TUNE")) {
        (void)tuning::parse(forced, t);
    }
    std::unique_ptr<plan, void (*)(plan *)> lead(build_plan(&cr, last->device, t, false, hb), release);
    std::unique_ptr<plan, void (*)(plan *)> p(new plan, release);
    p->cfg = *cfg;
    p->device = last->device;
    p->dbl = last->dbl;
    p->esize = last->esize;
    p->variants = last->variants;
    p->nvariants = last->nvariants;
    p->tune = last->tune;
    p->transform = rows * N;
    p->total = p->transform * cfg->batch_count;
    p->scale = norm_scale(cfg->normalization, cfg->direction, p->transform);
    p->real = cfg->direction == -1 ? 1 : 2;
    p->real_m = N / 2;
    device_guard guard(p->device);
    if (p->real == 2) {
        hip_check(hipMalloc(&p->real_scratch, cfg->batch_count * rows * hb * p->esize), "hipMalloc(real scratch)");
    }
    hip_check(hipEventCreateWithFlags(&p->last, hipEventDisableTiming | hipEventDisableSystemFence), "hipEventCreate");
    p->half = last.release();
    p->core = lead.release();
    trace(std::string(p->real == 1 ? "real r2c" : "real c2r") + " rank " + std::to_string(r) + " last axis n="
          + std::to_string(N));
    return p.release();
}

static std::string describe_plan(const plan &p)
{
    std::string s;
    if (p.layout) {
        const std::string c = describe_plan(*p.core);
        if (p.lay_fused) {
            return c;
        }
        return "layout-gather\n" + c.substr(0, c.rfind("tuning")) + "layout-scatter\n" + "tuning " + p.core->tune.str()
               + "\n";
    }
    if (p.real && p.core) {
        const std::string a = describe_plan(*p.half), b = describe_plan(*p.core);
        const std::string a0 = a.substr(0, a.rfind("tuning")), b0 = b.substr(0, b.rfind("tuning"));
        return (p.real == 1 ? a0 + b0 : b0 + a0) + "tuning " + p.half->tune.str() + "\n";
    }
    if (p.real && p.real_fused) {
        return p.launches[0].desc + "\ntuning " + p.half->tune.str() + "\n";
    }
    if (p.real) {
        const std::string pass = p.real == 1 ? "r2c-untangle" : "c2r-tangle";
        if (p.real == 2) {
            s += pass + " m=" + std::to_string(p.real_m) + "\n";
        }
        const std::string h = describe_plan(*p.half);
        s += h.substr(0, h.rfind("tuning"));
        if (p.real == 1) {
            s += pass + " m=" + std::to_string(p.real_m) + "\n";
        }
        s += "tuning " + p.half->tune.str() + "\n";
        return s;
    }
    if (p.copy_only) {
        s += "copy_scale\n";
    }
    for (const auto &l : p.launches) {
        s += l.desc;
        s += '\n';
    }
    s += "tuning " + p.tune.str() + "\n";
    return s;
}

} // namespace sfh

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------

using sfh::sfh_error;

extern "C" int syclfft_hip_supports(int device, char *err, size_t cap)
{
    try {
        std::string reason;
        if (sfh::check_device(device, reason) < 0) {
            sfh::set_err(err, cap, reason);
            return 0;
        }
        sfh::set_err(err, cap, "");
        return 1;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return 0;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return 0;
    }
}

extern "C" void *syclfft_hip_plan_create(const syclfft_hip_plan_config_v1 *config, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        return sfh::create_plan(config);
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
    } catch (const std::bad_alloc &) {
        sfh::set_err(err, cap, "Unable to allocate plan memory");
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
    }
    return nullptr;
}

extern "C" int syclfft_hip_plan_create_ex(
    const syclfft_hip_plan_config_v1 *config, void **plan, char *err, size_t cap)
{
    if (plan) {
        *plan = nullptr;
    }
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plan output pointer is null");
        }
        *plan = sfh::create_plan(config);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::bad_alloc &) {
        sfh::set_err(err, cap, "Unable to allocate plan memory");
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
    }
    return SYCLFFT_HIP_STATUS_PLANNING_FAILED;
}

extern "C" int syclfft_hip_real_plan_create(
    const syclfft_hip_plan_config_v1 *config, void **plan, char *err, size_t cap)
{
    if (plan) {
        *plan = nullptr;
    }
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plan output pointer is null");
        }
        *plan = sfh::create_real_plan(config);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::bad_alloc &) {
        sfh::set_err(err, cap, "Unable to allocate plan memory");
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
    }
    return SYCLFFT_HIP_STATUS_PLANNING_FAILED;
}

extern "C" int syclfft_hip_layout_plan_create(
    const syclfft_hip_plan_config_v1 *config, const syclfft_hip_layout_v1 *layout, void **plan, char *err,
    size_t cap)
{
    if (plan) {
        *plan = nullptr;
    }
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "plan output pointer is null");
        }
        *plan = sfh::create_layout_plan(config, layout);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::bad_alloc &) {
        sfh::set_err(err, cap, "Unable to allocate plan memory");
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
    }
    return SYCLFFT_HIP_STATUS_PLANNING_FAILED;
}

extern "C" int syclfft_hip_layout_execute(
    void *plan, const void *in, const void *in_imag, void *out, void *out_imag, void *stream, void *const *deps,
    size_t ndeps, void *done, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE, "plan handle is null");
        }
        auto &p = *static_cast<sfh::plan *>(plan);
        if (!p.layout) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "not a layout plan (syclfft_hip_layout_plan_create)");
        }
        if (ndeps && !deps) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "dependency list is null");
        }
        sfh::execute_layout(p, in, in_imag, out, out_imag, static_cast<hipStream_t>(stream), deps, ndeps,
                            static_cast<hipEvent_t>(done));
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

extern "C" void syclfft_hip_plan_destroy(void *plan)
{
    sfh::release(static_cast<sfh::plan *>(plan));
}

extern "C" uint64_t syclfft_hip_plan_scratch_bytes(const void *plan)
{
    if (!plan) {
        return 0;
    }
    const auto *p = static_cast<const sfh::plan *>(plan);
    if (p->real && p->core) {
        const uint64_t own = p->real == 2 ? p->total / (2 * p->real_m) * (p->real_m + 1) * p->esize : 0;
        return own + syclfft_hip_plan_scratch_bytes(p->half) + syclfft_hip_plan_scratch_bytes(p->core);
    }
    if (p->real) {
        return p->cfg.batch_count * p->real_m * p->esize + syclfft_hip_plan_scratch_bytes(p->half);
    }
    if (p->layout) {
        return (p->lay_fused ? 0 : p->total * p->esize) + syclfft_hip_plan_scratch_bytes(p->core);
    }
    return p->scratch_elems * p->esize;
}

extern "C" int syclfft_hip_plan_execute(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps,
    void *done, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE, "plan handle is null");
        }
        if (ndeps && !deps) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "dependency list is null");
        }
        sfh::execute_plan(
            *static_cast<sfh::plan *>(plan), in, out, static_cast<hipStream_t>(stream), deps, ndeps,
            static_cast<hipEvent_t>(done));
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

// Execution of a complex plan with `scratch` (device memory of `scratch_bytes` bytes on the
// plan's device) in place of the plan's own scratch; nothing is allocated.  Backs
// syclfft_hip_plan_execute_noalloc (fft_noalloc.hip), which owns the scratch memory.
extern "C" __attribute__((visibility("hidden"))) int sfh_plan_execute_with_scratch(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps, void *done,
    void *scratch, uint64_t scratch_bytes, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE, "plan handle is null");
        }
        if (ndeps && !deps) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "dependency list is null");
        }
        sfh::plan &p = *static_cast<sfh::plan *>(plan);
        if (p.real || p.layout) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                            "the no-allocation entry point runs complex plans only (not real or layout plans)");
        }
        const uint64_t need = p.scratch_elems * p.esize;
        if (need > scratch_bytes) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                            "plan needs " + std::to_string(need) + " bytes of scratch; the no-allocation entry point holds "
                                + std::to_string(scratch_bytes));
        }
        if (need && !scratch) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, "no scratch memory on the plan's device");
        }
        sfh::check_buffers(p, in, out);
        sfh::execute_checked(p, in, out, static_cast<hipStream_t>(stream), deps, ndeps, static_cast<hipEvent_t>(done),
                             scratch);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

// Device memory a complex plan holds: tables, scratch, arrival counters and its ring of
// pointer-array tables (0 for a null handle).
extern "C" __attribute__((visibility("hidden"))) uint64_t sfh_plan_device_bytes(const void *plan)
{
    if (!plan) {
        return 0;
    }
    const auto *p = static_cast<const sfh::plan *>(plan);
    uint64_t b = p->tables_bytes + p->scratch_elems * p->esize + p->pair_transforms * sizeof(uint32_t)
                 + static_cast<uint64_t>(sfh::plan::k_unit_slots) * p->unit_bytes;
    if (p->half) {
        b += sfh_plan_device_bytes(p->half);
    }
    if (p->core) {
        b += sfh_plan_device_bytes(p->core);
    }
    return b;
}

// Device of a plan (-1 for a null handle).
extern "C" __attribute__((visibility("hidden"))) int sfh_plan_device(const void *plan)
{
    return plan ? static_cast<const sfh::plan *>(plan)->device : -1;
}

extern "C" int syclfft_hip_plan_execute_many(
    void *plan, uint64_t count, const void *const *ins, void *const *outs, void *stream, void *const *deps,
    size_t ndeps, void *done, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        if (!plan) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_STATE, "plan handle is null");
        }
        if (ndeps && !deps) {
            throw sfh_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "dependency list is null");
        }
        sfh::execute_units(
            *static_cast<sfh::plan *>(plan), count, ins, outs, static_cast<hipStream_t>(stream), deps, ndeps,
            static_cast<hipEvent_t>(done));
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

// Same as syclfft_hip_plan_execute_many for callers inside the library whose buffers were
// validated as whole allocations already (the serving loop checks each request's buffers
// when it is submitted) and which order executions themselves (each request has its own
// buffers); only the per-transform placement rules are still checked, and executions of a
// plan without scratch may overlap.
extern "C" int syclfft_hip_execute_queue(
    void *const *plans, const void *const *ins, void *const *outs, size_t count, void *stream, void *done,
    char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        sfh::execute_queue(reinterpret_cast<sfh::plan *const *>(plans), ins, outs, count,
                           static_cast<hipStream_t>(stream), static_cast<hipEvent_t>(done));
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    } catch (...) {
        sfh::set_err(err, cap, "unknown error");
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

// Serving loop: class of `count` transforms of a plan for queue launches (-1: cannot
// join one; equal classes share a launch), and a queue execution of checked buffers.
extern "C" __attribute__((visibility("hidden"))) int sfh_plan_queue_class(const void *plan, uint64_t count)
{
    const sfh::plan &p = *static_cast<const sfh::plan *>(plan);
    if (!sfh::queueable_for(p, count)) {
        return -1;
    }
    return (p.dbl ? 1 << 16 : 0) + p.launches[0].variant;
}

extern "C" __attribute__((visibility("hidden"))) int sfh_execute_queue_checked_buffers(
    void *const *plans, const uint64_t *counts, const void *const *ins, void *const *outs, size_t n, void *stream,
    void *done, char *err, size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        sfh::execute_queue(reinterpret_cast<sfh::plan *const *>(plans), ins, outs, n, static_cast<hipStream_t>(stream),
                           static_cast<hipEvent_t>(done), counts, false, false);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

extern "C" __attribute__((visibility("hidden"))) int sfh_plan_execute_many_checked_buffers(
    void *plan, uint64_t count, const void *const *ins, void *const *outs, void *stream, void *done, char *err,
    size_t cap)
{
    try {
        sfh::set_err(err, cap, "");
        sfh::execute_units(*static_cast<sfh::plan *>(plan), count, ins, outs, static_cast<hipStream_t>(stream),
                           nullptr, 0, static_cast<hipEvent_t>(done), false, false);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const sfh_error &ex) {
        sfh::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

extern "C" size_t syclfft_hip_plan_describe(const void *plan, char *buf, size_t cap)
{
    if (!plan) {
        return 0;
    }
    const std::string s = sfh::describe_plan(*static_cast<const sfh::plan *>(plan));
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return s.size();
}

extern "C" uint32_t syclfft_hip_plan_launch_count(const void *plan)
{
    if (!plan) {
        return 0;
    }
    const auto *p = static_cast<const sfh::plan *>(plan);
    if (p->real && p->core) {
        return syclfft_hip_plan_launch_count(p->half) + syclfft_hip_plan_launch_count(p->core);
    }
    if (p->real) {
        return p->real_fused ? 1u : 1u + syclfft_hip_plan_launch_count(p->half);
    }
    if (p->layout) {
        return (p->lay_fused ? 0u : 2u) + syclfft_hip_plan_launch_count(p->core);
    }
    return static_cast<uint32_t>(p->launches.size()) + (p->copy_only ? 1u : 0u);
}

namespace
{
int api_supports(int device, char *err, size_t cap)
{
    return syclfft_hip_supports(device, err, cap);
}
void *api_create(const syclfft_hip_plan_config_v1 *config, char *err, size_t cap)
{
    return syclfft_hip_plan_create(config, err, cap);
}
void api_destroy(void *plan)
{
    syclfft_hip_plan_destroy(plan);
}
uint64_t api_scratch(const void *plan)
{
    return syclfft_hip_plan_scratch_bytes(plan);
}
int api_execute(
    void *plan, const void *in, void *out, void *stream, void *const *deps, size_t ndeps,
    void *done, char *err, size_t cap)
{
    return syclfft_hip_plan_execute(plan, in, out, stream, deps, ndeps, done, err, cap);
}
const syclfft_device_provider_v1 k_api = {
    SYCLFFT_DEVICE_PROVIDER_ABI_VERSION, "hip_mi355x", api_supports, api_create, api_destroy,
    api_scratch, api_execute};
} // namespace

extern "C" const syclfft_device_provider_v1 *syclfft_get_device_provider_v1(void)
{
    return &k_api;
}
