// Serving loop on the CPU emulation (TEST INFRASTRUCTURE ONLY): a mix of requests of
// several shapes, types, directions and normalizations is submitted at once (so waiting
// requests of one shape are merged into one launch), every result is checked against the
// FP64 oracle, then every request runs again alone and must give the same result within
// the oracle tolerance; also the budget and ticket error paths.  Prints "ok <check>" or
// "FAIL <check> ..." lines; exit status 1 on any failure.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "syclfft_hip_server.h"

extern "C" void oracle_fftn_ld(int rank, const uint64_t *lengths, uint64_t batch, const double *in,
                               double *out, int sign, int norm);

struct req_case
{
    std::vector<uint64_t> lens;
    uint64_t batch;
    bool dbl;
    int dir;
    uint32_t norm;
    bool in_place;
    std::vector<double> x;   // input (values as stored)
    void *din{}, *dout{};
    std::vector<double> mix; // result in the mix
};

static uint64_t prod(const std::vector<uint64_t> &v)
{
    uint64_t n = 1;
    for (auto x : v) {
        n *= x;
    }
    return n;
}

static void upload(req_case &c)
{
    const uint64_t total = prod(c.lens) * c.batch;
    const size_t es = c.dbl ? 16 : 8;
    hipMalloc(&c.din, total * es);
    if (c.in_place) {
        c.dout = c.din;
    } else {
        hipMalloc(&c.dout, total * es);
    }
    if (c.dbl) {
        std::memcpy(c.din, c.x.data(), total * 16);
    } else {
        float *f = static_cast<float *>(c.din);
        for (uint64_t i = 0; i < 2 * total; ++i) {
            f[i] = static_cast<float>(c.x[i]);
            c.x[i] = f[i];
        }
    }
}

static void reload(req_case &c)
{
    const uint64_t total = prod(c.lens) * c.batch;
    if (c.dbl) {
        std::memcpy(c.din, c.x.data(), total * 16);
    } else {
        float *f = static_cast<float *>(c.din);
        for (uint64_t i = 0; i < 2 * total; ++i) {
            f[i] = static_cast<float>(c.x[i]);
        }
    }
}

static std::vector<double> download(const req_case &c)
{
    const uint64_t total = prod(c.lens) * c.batch;
    std::vector<double> y(2 * total);
    if (c.dbl) {
        std::memcpy(y.data(), c.dout, total * 16);
    } else {
        const float *f = static_cast<const float *>(c.dout);
        for (uint64_t i = 0; i < 2 * total; ++i) {
            y[i] = f[i];
        }
    }
    return y;
}

static syclfft_hip_request_v1 make_req(const req_case &c, uint64_t tag)
{
    syclfft_hip_request_v1 r{};
    r.abi_version = SYCLFFT_HIP_SERVER_ABI_VERSION;
    r.scalar = c.dbl ? SYCLFFT_HIP_SCALAR_DOUBLE : SYCLFFT_HIP_SCALAR_FLOAT;
    r.direction = c.dir;
    r.rank = static_cast<uint32_t>(c.lens.size());
    for (size_t i = 0; i < c.lens.size(); ++i) {
        r.lengths[i] = c.lens[i];
    }
    r.batch_count = c.batch;
    r.normalization = c.norm;
    r.in_place = c.in_place ? 1u : 0u;
    r.in = c.din;
    r.out = c.dout;
    r.tag = tag;
    return r;
}

static double rel_l2(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

static double tol_of(const req_case &c)
{
    const double u = c.dbl ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
    return (c.lens.size() > 1 ? 12.0 : 4.0) * std::max(1.0, std::log2(static_cast<double>(prod(c.lens)))) * u;
}

int main()
{
    bool all = true;
    const auto check = [&](bool ok, const std::string &what, const std::string &detail = "") {
        std::printf("%s %s %s\n", ok ? "ok" : "FAIL", what.c_str(), detail.c_str());
        std::fflush(stdout);
        all = all && ok;
    };
    std::mt19937_64 rng(20260005);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const std::vector<std::vector<uint64_t>> shapes = {{64}, {300}, {4096}, {1009}, {13, 16}, {8, 8, 4}, {32768}, {10007}, {2520}};
    std::vector<req_case> cases;
    for (int k = 0; k < 36; ++k) {
        req_case c;
        c.lens = shapes[k % shapes.size()];
        c.batch = 1 + static_cast<uint64_t>(rng() % 4);
        c.dbl = (k % 5) == 4;
        c.dir = (k / 12) % 2 ? 1 : -1; // k and k + 6 often share a shape key: they merge
        c.norm = static_cast<uint32_t>((k / 24) % 4 + (k / 12) % 2);
        c.in_place = (k % 7) == 3;
        c.x.resize(2 * prod(c.lens) * c.batch);
        for (auto &v : c.x) {
            v = u(rng);
        }
        upload(c);
        cases.push_back(std::move(c));
    }
    char err[1024];
    void *srv = syclfft_hip_server_create(-1, uint64_t(1) << 30, 2, 64, err, sizeof(err));
    check(srv != nullptr, "server_create", err);
    if (!srv) {
        return 1;
    }
    // the mix: everything submitted at once
    std::vector<uint64_t> tickets;
    for (size_t k = 0; k < cases.size(); ++k) {
        const syclfft_hip_request_v1 r = make_req(cases[k], k);
        uint64_t t = 0;
        const int st = syclfft_hip_server_submit(srv, &r, &t, err, sizeof(err));
        check(st == 0, "submit " + std::to_string(k), err);
        tickets.push_back(t);
    }
    uint32_t max_group = 0;
    for (size_t k = 0; k < cases.size(); ++k) {
        syclfft_hip_completion_v1 c{};
        const int st = syclfft_hip_server_wait(srv, tickets[k], &c, err, sizeof(err));
        check(st == 0 && c.tag == k && c.latency_seconds >= 0, "wait " + std::to_string(k), err);
        max_group = std::max(max_group, c.group_size);
        cases[k].mix = download(cases[k]);
    }
    syclfft_hip_server_stats_v1 s{};
    syclfft_hip_server_stats(srv, &s);
    check(s.completed == cases.size() && s.failed == 0, "stats completed",
          std::to_string(s.completed) + "/" + std::to_string(cases.size()));
    check(s.launches < cases.size() && max_group > 1, "requests merged into shared launches",
          "launches=" + std::to_string(s.launches) + " max_group=" + std::to_string(max_group));
    check(s.peak_footprint_bytes <= s.budget_bytes, "footprint within budget");
    // oracle, then each request alone
    for (size_t k = 0; k < cases.size(); ++k) {
        req_case &c = cases[k];
        std::vector<double> ref(c.x.size());
        oracle_fftn_ld(static_cast<int>(c.lens.size()), c.lens.data(), c.batch, c.x.data(), ref.data(), c.dir,
                       static_cast<int>(c.norm));
        const double e = rel_l2(c.mix, ref);
        char d[128];
        std::snprintf(d, sizeof(d), "relL2=%.3e tol=%.1e", e, tol_of(c));
        check(e <= tol_of(c), "mix vs oracle " + std::to_string(k), d);
        reload(c);
        const syclfft_hip_request_v1 r = make_req(c, 1000 + k);
        uint64_t t = 0;
        syclfft_hip_completion_v1 comp{};
        const int st = syclfft_hip_server_submit(srv, &r, &t, err, sizeof(err));
        const int wt = st == 0 ? syclfft_hip_server_wait(srv, t, &comp, err, sizeof(err)) : st;
        const std::vector<double> alone = download(c);
        const double ea = rel_l2(alone, c.mix);
        std::snprintf(d, sizeof(d), "relL2(alone, mix)=%.3e group=%u", ea, comp.group_size);
        check(wt == 0 && comp.group_size == 1 && ea <= tol_of(c), "alone vs mix " + std::to_string(k), d);
    }
    // the mix again with every plan resident: waiting requests of other shapes whose plans
    // share a kernel variant now merge into work-queue launches; results as in the first mix
    {
        for (auto &c : cases) {
            reload(c);
        }
        std::vector<uint64_t> t2;
        for (size_t k = 0; k < cases.size(); ++k) {
            const syclfft_hip_request_v1 r = make_req(cases[k], 2000 + k);
            uint64_t t = 0;
            const int st = syclfft_hip_server_submit(srv, &r, &t, err, sizeof(err));
            check(st == 0, "resubmit " + std::to_string(k), err);
            t2.push_back(t);
        }
        uint32_t g2 = 0;
        for (size_t k = 0; k < cases.size(); ++k) {
            syclfft_hip_completion_v1 c{};
            const int st = syclfft_hip_server_wait(srv, t2[k], &c, err, sizeof(err));
            g2 = std::max(g2, c.group_size);
            const std::vector<double> again = download(cases[k]);
            const double e = rel_l2(again, cases[k].mix);
            char d[128];
            std::snprintf(d, sizeof(d), "relL2(second mix, first mix)=%.3e group=%u", e, c.group_size);
            check(st == 0 && e <= tol_of(cases[k]), "second mix " + std::to_string(k), d);
        }
        check(g2 > 1, "second mix merged requests", "max_group=" + std::to_string(g2));
    }
    // error paths
    {
        req_case big;
        big.lens = {1 << 12};
        big.batch = 64;
        big.dbl = false;
        big.dir = -1;
        big.norm = 0;
        big.in_place = false;
        big.x.assign(2 * prod(big.lens) * big.batch, 0.0);
        upload(big);
        void *small = syclfft_hip_server_create(-1, 1 << 20, 1, 0, err, sizeof(err));
        const syclfft_hip_request_v1 r = make_req(big, 7);
        uint64_t t = 0;
        const int st = syclfft_hip_server_submit(small, &r, &t, err, sizeof(err));
        check(st == SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "footprint above budget rejected", err);
        syclfft_hip_completion_v1 c{};
        const int wt = syclfft_hip_server_wait(small, 12345, &c, err, sizeof(err));
        check(wt == SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "unknown ticket rejected", err);
        std::vector<double> host(16);
        syclfft_hip_request_v1 h = make_req(big, 8);
        h.in = host.data();
        const int hs = syclfft_hip_server_submit(small, &h, &t, err, sizeof(err));
        check(hs == SYCLFFT_HIP_STATUS_INVALID_POINTER, "host buffer rejected", err);
        syclfft_hip_server_destroy(small);
        hipFree(big.din);
        hipFree(big.dout);
    }
    {
        // Eviction under a tight budget: three 256-point plans (different keys) fill it, then
        // a request of the first, least recently used shape arrives.  Its resident plan is the
        // one to use, and it must survive the eviction that makes room for the request's
        // buffers (another plan goes, or the request waits).
        std::vector<req_case> ev;
        const int dirs[3] = {-1, 1, -1};
        const uint32_t norms[3] = {0, 0, 2};
        for (int k = 0; k < 3; ++k) {
            req_case c;
            c.lens = {256};
            c.batch = k == 0 ? 2 : 1; // the first shape's requests need the most room
            c.dbl = false;
            c.dir = dirs[k];
            c.norm = norms[k];
            c.in_place = false;
            c.x.resize(2 * 256 * c.batch);
            for (auto &v : c.x) {
                v = std::uniform_real_distribution<double>(-1, 1)(rng);
            }
            upload(c);
            ev.push_back(std::move(c));
        }
        // A plan is charged min(its admission bound (2 * 8 + 4) * 256 * 8 B = 40960 B, what it
        // holds + 64 KiB) = 40960 B.  Two plans, a new one and a 4 KiB request fit; three
        // plans and the first shape's 8 KiB request do not.
        void *tight = syclfft_hip_server_create(-1, 128000, 1, 8, err, sizeof(err));
        check(tight != nullptr, "tight server_create", err);
        const int order[6] = {0, 1, 2, 0, 1, 2};
        for (int k = 0; k < 6; ++k) {
            req_case &c = ev[order[k]];
            const syclfft_hip_request_v1 r = make_req(c, 100 + k);
            uint64_t t = 0;
            int st = syclfft_hip_server_submit(tight, &r, &t, err, sizeof(err));
            syclfft_hip_completion_v1 comp{};
            if (st == SYCLFFT_HIP_STATUS_OK) {
                st = syclfft_hip_server_wait(tight, t, &comp, err, sizeof(err));
            }
            check(st == SYCLFFT_HIP_STATUS_OK, "tight budget request " + std::to_string(k), err);
            const uint64_t total = 256 * c.batch;
            const std::vector<double> got = download(c);
            std::vector<double> want(2 * total);
            const uint64_t lens[3] = {256, 1, 1};
            oracle_fftn_ld(1, lens, c.batch, c.x.data(), want.data(), c.dir, static_cast<int>(c.norm));
            const double e = rel_l2(got, want);
            check(e < tol_of(c), "tight budget result " + std::to_string(k));
        }
        syclfft_hip_server_stats_v1 s{};
        syclfft_hip_server_stats(tight, &s);
        check(s.plans_evicted > 0, "plans evicted under the tight budget");
        check(s.peak_footprint_bytes <= s.budget_bytes, "tight footprint within budget");
        syclfft_hip_server_destroy(tight);
        for (auto &c : ev) {
            hipFree(c.din);
            hipFree(c.dout);
        }
    }
    syclfft_hip_server_destroy(srv);
    for (auto &c : cases) {
        hipFree(c.din);
        if (!c.in_place) {
            hipFree(c.dout);
        }
    }
    return all ? 0 : 1;
}
