// Runs provider plans on the CPU emulation and checks them against the FP64 oracle
// (TEST INFRASTRUCTURE ONLY).  Usage: emu_driver <case>...  with a case written as
//   f|d|b|h:n0[xn1[xn2]]:batch:dir:norm:inplace      e.g.  f:4096:2:-1:0:0
// (b, h: float plans with split-precision BF16 / FP16 MFMA radix-16 passes)
// Prints one line per case: "ok <case> relL2=<e>" or "FAIL <case> ..."; exit status 1 on
// any failure.  EMU_MANY=1 executes through syclfft_hip_plan_execute_many with every
// transform in its own allocation (allocated in reverse order); EMU_MANY=2 does the same for
// the first half of the batch only (a count below the plan's batch count); EMU_MANY=3 passes
// pointers into the contiguous batch for its first half (no unit tables).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "syclfft_hip_provider.h"

extern "C" void oracle_fftn_ld(int rank, const uint64_t *lengths, uint64_t batch, const double *in,
                               double *out, int sign, int norm);

// Real plans ('R' float, 'Q' double), R:n0[xn1[xn2]]:batch:dir:norm:0 with the last length N
// even: direction -1 = R2C (real [batch][..][N] -> complex [batch][..][N/2 + 1]), +1 = C2R
// (a Hermitian half spectrum -> real); checked against the complex FP64 oracle.
static bool run_real(const std::string &spec)
{
    char prec = 0, dims[128] = {0};
    unsigned long long batch = 1;
    int dir = -1, norm = 0, inplace = 0;
    if (std::sscanf(spec.c_str(), "%c:%127[0-9x]:%llu:%d:%d:%d", &prec, dims, &batch, &dir, &norm, &inplace) != 6) {
        std::printf("FAIL %s bad case spec\n", spec.c_str());
        return false;
    }
    const bool dbl = prec == 'Q';
    syclfft_hip_plan_config_v1 cfg{};
    cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
    cfg.scalar = dbl ? SYCLFFT_HIP_SCALAR_DOUBLE : SYCLFFT_HIP_SCALAR_FLOAT;
    cfg.direction = dir;
    uint64_t lens[3] = {1, 1, 1}, T = 1;
    {
        std::string d(dims);
        size_t pos = 0;
        for (;;) {
            const size_t x = d.find('x', pos);
            lens[cfg.rank] = std::strtoull(d.substr(pos, x == std::string::npos ? std::string::npos : x - pos).c_str(),
                                           nullptr, 10);
            cfg.lengths[cfg.rank] = lens[cfg.rank];
            T *= lens[cfg.rank];
            ++cfg.rank;
            if (x == std::string::npos || cfg.rank == 3) {
                break;
            }
            pos = x + 1;
        }
    }
    cfg.batch_count = batch;
    cfg.normalization = static_cast<uint32_t>(norm);
    cfg.device = -1;
    const uint64_t n = lens[cfg.rank - 1], hb = n / 2 + 1, rows = T / n;
    char err[1024];
    void *plan = nullptr;
    if (syclfft_hip_real_plan_create(&cfg, &plan, err, sizeof(err)) != 0) {
        std::printf("FAIL %s create: %s\n", spec.c_str(), err);
        return false;
    }
    std::mt19937_64 rng(20261100 + T + batch);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    // a real signal (batch x T) and its full spectrum (FP64 oracle, no normalization)
    std::vector<double> rc(2 * T * batch, 0.0), spec_full(2 * T * batch);
    for (uint64_t i = 0; i < T * batch; ++i) {
        const double v = u(rng);
        rc[2 * i] = dbl ? v : static_cast<double>(static_cast<float>(v));
    }
    oracle_fftn_ld(static_cast<int>(cfg.rank), lens, batch, rc.data(), spec_full.data(), -1, 0);
    const size_t rs = dbl ? 8 : 4; // bytes per real value
    const uint64_t half_vals = 2 * rows * hb * batch;
    const uint64_t in_vals = dir == -1 ? T * batch : half_vals;
    const uint64_t out_vals = dir == -1 ? half_vals : T * batch;
    void *din = nullptr, *dout = nullptr;
    hipMalloc(&din, in_vals * rs);
    hipMalloc(&dout, out_vals * rs);
    std::vector<double> in(in_vals), expect(out_vals);
    // flat index of the element mirrored through every axis (the Hermitian partner)
    const auto mirror = [&](uint64_t e) {
        uint64_t m = 0, rest = e, mul = 1;
        std::vector<uint64_t> idx(cfg.rank);
        for (int d = static_cast<int>(cfg.rank) - 1; d >= 0; --d) {
            idx[d] = rest % lens[d];
            rest /= lens[d];
        }
        for (int d = static_cast<int>(cfg.rank) - 1; d >= 0; --d) {
            m += ((lens[d] - idx[d]) % lens[d]) * mul;
            mul *= lens[d];
        }
        return m;
    };
    if (dir == -1) {
        for (uint64_t i = 0; i < T * batch; ++i) {
            in[i] = rc[2 * i];
        }
        std::vector<double> ref(2 * T * batch);
        oracle_fftn_ld(static_cast<int>(cfg.rank), lens, batch, rc.data(), ref.data(), -1, norm);
        for (uint64_t b = 0; b < batch; ++b) {
            for (uint64_t r = 0; r < rows; ++r) {
                for (uint64_t k = 0; k < hb; ++k) {
                    const uint64_t o = (b * rows + r) * hb + k, f = b * T + r * n + k;
                    expect[2 * o] = ref[2 * f];
                    expect[2 * o + 1] = ref[2 * f + 1];
                }
            }
        }
    } else {
        // the input: the half spectrum, made exactly Hermitian (S[e] = conj S[mirror e]) and
        // rounded to the plan's type; the expected output: the oracle's inverse of the full
        // spectrum that half defines
        std::vector<double> full(2 * T * batch), ref(2 * T * batch);
        for (uint64_t b = 0; b < batch; ++b) {
            const double *S = spec_full.data() + 2 * b * T;
            double *F = full.data() + 2 * b * T;
            for (uint64_t e = 0; e < T; ++e) {
                const uint64_t m = mirror(e);
                double re = 0.5 * (S[2 * e] + S[2 * m]), im = 0.5 * (S[2 * e + 1] - S[2 * m + 1]);
                if (!dbl) {
                    re = static_cast<float>(re);
                    im = static_cast<float>(im);
                }
                F[2 * e] = re;
                F[2 * e + 1] = im;
            }
            for (uint64_t r = 0; r < rows; ++r) {
                for (uint64_t k = 0; k < hb; ++k) {
                    const uint64_t o = (b * rows + r) * hb + k, f = r * n + k;
                    in[2 * o] = F[2 * f];
                    in[2 * o + 1] = F[2 * f + 1];
                }
            }
        }
        oracle_fftn_ld(static_cast<int>(cfg.rank), lens, batch, full.data(), ref.data(), 1, norm);
        for (uint64_t i = 0; i < T * batch; ++i) {
            expect[i] = ref[2 * i];
        }
    }
    if (dbl) {
        std::memcpy(din, in.data(), in_vals * 8);
    } else {
        float *f = static_cast<float *>(din);
        for (uint64_t i = 0; i < in_vals; ++i) {
            f[i] = static_cast<float>(in[i]);
        }
    }
    if (syclfft_hip_plan_execute(plan, din, dout, nullptr, nullptr, 0, nullptr, err, sizeof(err)) != 0) {
        std::printf("FAIL %s execute: %s\n", spec.c_str(), err);
        return false;
    }
    double num = 0, den = 0;
    for (uint64_t i = 0; i < out_vals; ++i) {
        const double g = dbl ? static_cast<double *>(dout)[i] : static_cast<float *>(dout)[i];
        num += (g - expect[i]) * (g - expect[i]);
        den += expect[i] * expect[i];
    }
    const double rel = std::sqrt(num / (den > 0 ? den : 1));
    const double u_ = dbl ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
    const double tol = 4.0 * std::max(1.0, std::log2(static_cast<double>(T))) * u_;
    char desc[4096];
    syclfft_hip_plan_describe(plan, desc, sizeof(desc));
    for (char *c = desc; *c; ++c) {
        if (*c == '\n') {
            *c = '|';
        }
    }
    const bool ok = rel <= tol && std::isfinite(rel);
    std::printf("%s %s relL2=%.3e tol=%.1e launches=%u [%s]\n", ok ? "ok" : "FAIL", spec.c_str(), rel, tol,
                syclfft_hip_plan_launch_count(plan), desc);
    std::fflush(stdout);
    syclfft_hip_plan_destroy(plan);
    hipFree(din);
    hipFree(dout);
    return ok;
}

// Layout plans ('L' float, 'M' double):
//   L:n0[xn1[xn2]]:batch:dir:norm:inplace:istride:idist:ostride:odist:iplanar:oplanar
// The input layout's gaps hold a sentinel that must not leak into the result; the output
// layout's gaps (out of place) must keep theirs.
static bool run_layout(const std::string &spec)
{
    char prec = 0, dims[128] = {0};
    unsigned long long batch = 1, is = 1, id = 1, os = 1, od = 1;
    int dir = -1, norm = 0, inplace = 0, ip = 0, op = 0;
    if (std::sscanf(spec.c_str(), "%c:%127[0-9x]:%llu:%d:%d:%d:%llu:%llu:%llu:%llu:%d:%d", &prec, dims, &batch, &dir,
                    &norm, &inplace, &is, &id, &os, &od, &ip, &op) != 12) {
        std::printf("FAIL %s bad case spec\n", spec.c_str());
        return false;
    }
    const bool dbl = prec == 'M';
    syclfft_hip_plan_config_v1 cfg{};
    cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
    cfg.scalar = dbl ? SYCLFFT_HIP_SCALAR_DOUBLE : SYCLFFT_HIP_SCALAR_FLOAT;
    cfg.direction = dir;
    uint64_t lens[3] = {1, 1, 1}, T = 1;
    {
        std::string d(dims);
        size_t pos = 0;
        while (cfg.rank < 3 && pos <= d.size()) {
            const size_t x = d.find('x', pos);
            lens[cfg.rank] = std::strtoull(d.substr(pos, x == std::string::npos ? std::string::npos : x - pos).c_str(),
                                           nullptr, 10);
            T *= lens[cfg.rank];
            cfg.lengths[cfg.rank] = lens[cfg.rank];
            ++cfg.rank;
            if (x == std::string::npos) {
                break;
            }
            pos = x + 1;
        }
    }
    cfg.batch_count = batch;
    cfg.in_place = static_cast<uint32_t>(inplace);
    cfg.normalization = static_cast<uint32_t>(norm);
    cfg.device = -1;
    syclfft_hip_layout_v1 lay{is, id, os, od, static_cast<uint32_t>(ip), static_cast<uint32_t>(op)};
    char err[1024];
    void *plan = nullptr;
    if (syclfft_hip_layout_plan_create(&cfg, &lay, &plan, err, sizeof(err)) != 0) {
        std::printf("FAIL %s create: %s\n", spec.c_str(), err);
        return false;
    }
    const double sentinel = 4096.0;
    const uint64_t iext = (batch - 1) * id + (T - 1) * is + 1, oext = (batch - 1) * od + (T - 1) * os + 1;
    // host images of each array in doubles: interleaved sides 2 per element, planar sides 1 per array
    std::vector<double> in_re(ip ? iext : 2 * iext, sentinel), in_im(ip ? iext : 0, sentinel);
    std::vector<double> x(2 * T * batch), ref(2 * T * batch);
    std::mt19937_64 rng(20261300 + T + batch);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (uint64_t b = 0; b < batch; ++b) {
        for (uint64_t e = 0; e < T; ++e) {
            double re = u(rng), im = u(rng);
            if (!dbl) {
                re = static_cast<float>(re);
                im = static_cast<float>(im);
            }
            x[2 * (b * T + e)] = re;
            x[2 * (b * T + e) + 1] = im;
            const uint64_t o = b * id + e * is;
            if (ip) {
                in_re[o] = re;
                in_im[o] = im;
            } else {
                in_re[2 * o] = re;
                in_re[2 * o + 1] = im;
            }
        }
    }
    oracle_fftn_ld(static_cast<int>(cfg.rank), lens, batch, x.data(), ref.data(), dir, norm);
    const size_t rs = dbl ? 8 : 4;
    const auto upload = [&](const std::vector<double> &h) -> void * {
        void *d = nullptr;
        if (h.empty()) {
            return nullptr;
        }
        hipMalloc(&d, h.size() * rs);
        for (size_t i = 0; i < h.size(); ++i) {
            if (dbl) {
                static_cast<double *>(d)[i] = h[i];
            } else {
                static_cast<float *>(d)[i] = static_cast<float>(h[i]);
            }
        }
        return d;
    };
    void *d_in = upload(in_re), *d_in_im = upload(in_im);
    void *d_out = d_in, *d_out_im = d_in_im;
    if (!inplace) {
        d_out = upload(std::vector<double>(op ? oext : 2 * oext, sentinel));
        d_out_im = op ? upload(std::vector<double>(oext, sentinel)) : nullptr;
    }
    if (syclfft_hip_layout_execute(plan, d_in, d_in_im, d_out, d_out_im, nullptr, nullptr, 0, nullptr, err,
                                   sizeof(err))
        != 0) {
        std::printf("FAIL %s execute: %s\n", spec.c_str(), err);
        return false;
    }
    const auto val = [&](void *d, uint64_t i) {
        return dbl ? static_cast<double *>(d)[i] : static_cast<double>(static_cast<float *>(d)[i]);
    };
    double num = 0, den = 0;
    std::vector<char> touched(op ? oext : 2 * oext, 0);
    for (uint64_t b = 0; b < batch; ++b) {
        for (uint64_t e = 0; e < T; ++e) {
            const uint64_t o = b * od + e * os;
            const double gr = op ? val(d_out, o) : val(d_out, 2 * o);
            const double gi = op ? val(d_out_im, o) : val(d_out, 2 * o + 1);
            const double er = ref[2 * (b * T + e)], ei = ref[2 * (b * T + e) + 1];
            num += (gr - er) * (gr - er) + (gi - ei) * (gi - ei);
            den += er * er + ei * ei;
            if (op) {
                touched[o] = 1;
            } else {
                touched[2 * o] = touched[2 * o + 1] = 1;
            }
        }
    }
    bool gaps_ok = true;
    if (!inplace) {
        for (size_t i = 0; i < touched.size(); ++i) {
            if (!touched[i] && (val(d_out, i) != sentinel || (op && val(d_out_im, i) != sentinel))) {
                gaps_ok = false;
            }
        }
    }
    const double rel = std::sqrt(num / (den > 0 ? den : 1));
    const double u_ = dbl ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
    const double tol = 4.0 * std::max(1.0, std::log2(static_cast<double>(T))) * u_;
    char desc[2048];
    syclfft_hip_plan_describe(plan, desc, sizeof(desc));
    for (char *c = desc; *c; ++c) {
        if (*c == '\n') {
            *c = '|';
        }
    }
    const bool ok = rel <= tol && std::isfinite(rel) && gaps_ok;
    std::printf("%s %s relL2=%.3e tol=%.1e gaps=%s launches=%u [%s]\n", ok ? "ok" : "FAIL", spec.c_str(), rel, tol,
                gaps_ok ? "kept" : "WRITTEN", syclfft_hip_plan_launch_count(plan), desc);
    std::fflush(stdout);
    syclfft_hip_plan_destroy(plan);
    for (void *d : {d_in, d_in_im}) {
        if (d) {
            hipFree(d);
        }
    }
    if (!inplace) {
        for (void *d : {d_out, d_out_im}) {
            if (d) {
                hipFree(d);
            }
        }
    }
    return ok;
}

static bool run_case(const std::string &spec)
{
    if (!spec.empty() && (spec[0] == 'L' || spec[0] == 'M')) {
        return run_layout(spec);
    }
    if (!spec.empty() && (spec[0] == 'R' || spec[0] == 'Q')) {
        return run_real(spec);
    }
    char prec = 0;
    char dims[128] = {0};
    unsigned long long batch = 1;
    int dir = -1, norm = 0, inplace = 0;
    if (std::sscanf(spec.c_str(), "%c:%127[0-9x]:%llu:%d:%d:%d", &prec, dims, &batch, &dir, &norm,
                    &inplace) != 6) {
        std::printf("FAIL %s bad case spec\n", spec.c_str());
        return false;
    }
    std::vector<uint64_t> lens;
    for (char *tok = std::strtok(dims, "x"); tok; tok = std::strtok(nullptr, "x")) {
        lens.push_back(std::strtoull(tok, nullptr, 10));
    }
    // 'f' float, 'd' double, 'b' / 'h': float with split-precision MFMA radix-16 passes
    // (BF16 / scaled FP16 hi+lo operands)
    const bool dbl = prec == 'd';
    const uint32_t split = prec == 'b' ? SYCLFFT_HIP_PRECISION_SPLIT_BF16
                         : prec == 'h' ? SYCLFFT_HIP_PRECISION_SPLIT_FP16
                                       : SYCLFFT_HIP_PRECISION_VALU;
    syclfft_hip_plan_config_v1 cfg{};
    cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
    cfg.scalar = dbl ? SYCLFFT_HIP_SCALAR_DOUBLE : SYCLFFT_HIP_SCALAR_FLOAT;
    cfg.direction = dir;
    cfg.rank = static_cast<uint32_t>(lens.size());
    for (size_t i = 0; i < lens.size(); ++i) {
        cfg.lengths[i] = lens[i];
    }
    cfg.batch_count = batch;
    cfg.in_place = static_cast<uint32_t>(inplace);
    cfg.normalization = static_cast<uint32_t>(norm);
    cfg.device = -1;
    cfg.measure = std::getenv("EMU_MEASURE") ? 1u : 0u;
    cfg.precision = split;
    char err[1024];
    void *plan = nullptr;
    const int st = syclfft_hip_plan_create_ex(&cfg, &plan, err, sizeof(err));
    if (st != 0) {
        std::printf("FAIL %s create status %d: %s\n", spec.c_str(), st, err);
        return false;
    }
    uint64_t total = batch;
    for (auto n : lens) {
        total *= n;
    }
    std::vector<double> x(2 * total), ref(2 * total), got(2 * total);
    std::mt19937_64 rng(20260000 + total);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto &v : x) {
        v = u(rng);
    }
    const size_t es = dbl ? 16 : 8;
    void *din = nullptr, *dout = nullptr;
    hipMalloc(&din, total * es);
    if (inplace) {
        dout = din;
    } else {
        hipMalloc(&dout, total * es);
    }
    if (dbl) {
        std::memcpy(din, x.data(), total * 16);
    } else {
        float *f = static_cast<float *>(din);
        for (uint64_t i = 0; i < 2 * total; ++i) {
            f[i] = static_cast<float>(x[i]);
            x[i] = f[i];
        }
    }
    const char *many_env = std::getenv("EMU_MANY");
    const int many = many_env ? std::atoi(many_env) : 0;
    uint64_t tsz = 1;
    for (auto n : lens) {
        tsz *= n;
    }
    uint64_t count = batch;
    if (many == 3) {
        // pointers into the contiguous batch (the affine case: no unit tables), first half
        count = (batch + 1) / 2;
        std::vector<const void *> ins(count);
        std::vector<void *> outs(count);
        for (uint64_t k = 0; k < count; ++k) {
            ins[k] = static_cast<unsigned char *>(din) + k * tsz * es;
            outs[k] = static_cast<unsigned char *>(dout) + k * tsz * es;
        }
        const int est = syclfft_hip_plan_execute_many(plan, count, ins.data(), outs.data(), nullptr, nullptr, 0,
                                                      nullptr, err, sizeof(err));
        if (est != 0) {
            std::printf("FAIL %s execute_many status %d: %s\n", spec.c_str(), est, err);
            return false;
        }
    } else if (many) {
        count = many == 2 ? (batch + 1) / 2 : batch;
        std::vector<void *> ins(count), outs(count);
        for (uint64_t k = count; k-- > 0;) {
            hipMalloc(&ins[k], tsz * es);
            std::memcpy(ins[k], static_cast<unsigned char *>(din) + k * tsz * es, tsz * es);
            if (inplace) {
                outs[k] = ins[k];
            } else {
                hipMalloc(&outs[k], tsz * es);
            }
        }
        const int est = syclfft_hip_plan_execute_many(plan, count, ins.data(), outs.data(), nullptr, nullptr, 0,
                                                      nullptr, err, sizeof(err));
        if (est != 0) {
            std::printf("FAIL %s execute_many status %d: %s\n", spec.c_str(), est, err);
            return false;
        }
        for (uint64_t k = 0; k < count; ++k) {
            std::memcpy(static_cast<unsigned char *>(dout) + k * tsz * es, outs[k], tsz * es);
            if (!inplace) {
                hipFree(outs[k]);
            }
            hipFree(ins[k]);
        }
    } else {
        const int est =
            syclfft_hip_plan_execute(plan, din, dout, nullptr, nullptr, 0, nullptr, err, sizeof(err));
        if (est != 0) {
            std::printf("FAIL %s execute status %d: %s\n", spec.c_str(), est, err);
            return false;
        }
    }
    total = count * tsz; // compare the transforms that ran
    if (dbl) {
        std::memcpy(got.data(), dout, total * 16);
    } else {
        const float *f = static_cast<const float *>(dout);
        for (uint64_t i = 0; i < 2 * total; ++i) {
            got[i] = f[i];
        }
    }
    oracle_fftn_ld(static_cast<int>(lens.size()), lens.data(), count, x.data(), ref.data(), dir, norm);
    double num = 0, den = 0;
    for (uint64_t i = 0; i < 2 * total; ++i) {
        num += (got[i] - ref[i]) * (got[i] - ref[i]);
        den += ref[i] * ref[i];
    }
    const double rel = std::sqrt(num / (den > 0 ? den : 1));
    uint64_t tsize = 1;
    for (auto n : lens) {
        tsize *= n;
    }
    // unit roundoff of the arithmetic: split BF16 operands carry 16 bits (hi + lo, lo*lo
    // dropped), scaled split FP16 operands 22 bits
    const double u_ = dbl ? std::ldexp(1.0, -53)
                    : split == SYCLFFT_HIP_PRECISION_SPLIT_BF16 ? std::ldexp(1.0, -16)
                    : split == SYCLFFT_HIP_PRECISION_SPLIT_FP16 ? std::ldexp(1.0, -22)
                                                                 : std::ldexp(1.0, -24);
    const double tol = 4.0 * std::max(1.0, std::log2(static_cast<double>(tsize))) * u_;
    char desc[4096];
    syclfft_hip_plan_describe(plan, desc, sizeof(desc));
    for (char *c = desc; *c; ++c) {
        if (*c == '\n') {
            *c = '|';
        }
    }
    const bool ok = rel <= tol && std::isfinite(rel);
    std::printf("%s %s relL2=%.3e tol=%.1e launches=%u [%s]\n", ok ? "ok" : "FAIL", spec.c_str(), rel,
                tol, syclfft_hip_plan_launch_count(plan), desc);
    std::fflush(stdout);
    syclfft_hip_plan_destroy(plan);
    hipFree(din);
    if (!inplace) {
        hipFree(dout);
    }
    return ok;
}

// EMU_QUEUE=1: every case's plan is created first and all executions go through one
// syclfft_hip_execute_queue call (plain single-launch plans share work-queue launches).
struct queued
{
    std::string spec;
    std::vector<uint64_t> lens;
    uint64_t batch{}, total{};
    int dir{}, norm{}, inplace{};
    bool dbl{};
    void *plan{}, *din{}, *dout{};
    std::vector<double> x;
};

static bool run_queue(int argc, char **argv)
{
    std::vector<queued> qs;
    char err[1024];
    for (int i = 1; i < argc; ++i) {
        queued q;
        q.spec = argv[i];
        char prec = 0;
        char dims[128] = {0};
        unsigned long long batch = 1;
        if (std::sscanf(argv[i], "%c:%127[0-9x]:%llu:%d:%d:%d", &prec, dims, &batch, &q.dir, &q.norm,
                        &q.inplace) != 6) {
            std::printf("FAIL %s bad case spec\n", argv[i]);
            return false;
        }
        for (char *tok = std::strtok(dims, "x"); tok; tok = std::strtok(nullptr, "x")) {
            q.lens.push_back(std::strtoull(tok, nullptr, 10));
        }
        q.dbl = prec == 'd';
        q.batch = batch;
        syclfft_hip_plan_config_v1 cfg{};
        cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
        cfg.scalar = q.dbl ? SYCLFFT_HIP_SCALAR_DOUBLE : SYCLFFT_HIP_SCALAR_FLOAT;
        cfg.direction = q.dir;
        cfg.rank = static_cast<uint32_t>(q.lens.size());
        for (size_t k = 0; k < q.lens.size(); ++k) {
            cfg.lengths[k] = q.lens[k];
        }
        cfg.batch_count = batch;
        cfg.in_place = static_cast<uint32_t>(q.inplace);
        cfg.normalization = static_cast<uint32_t>(q.norm);
        cfg.device = -1;
        if (syclfft_hip_plan_create_ex(&cfg, &q.plan, err, sizeof(err)) != 0) {
            std::printf("FAIL %s create: %s\n", argv[i], err);
            return false;
        }
        q.total = batch;
        for (auto n : q.lens) {
            q.total *= n;
        }
        q.x.resize(2 * q.total);
        std::mt19937_64 rng(20260000 + q.total + i);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        for (auto &v : q.x) {
            v = u(rng);
        }
        const size_t es = q.dbl ? 16 : 8;
        hipMalloc(&q.din, q.total * es);
        q.dout = q.din;
        if (!q.inplace) {
            hipMalloc(&q.dout, q.total * es);
        }
        if (q.dbl) {
            std::memcpy(q.din, q.x.data(), q.total * 16);
        } else {
            float *f = static_cast<float *>(q.din);
            for (uint64_t k = 0; k < 2 * q.total; ++k) {
                f[k] = static_cast<float>(q.x[k]);
                q.x[k] = f[k];
            }
        }
        qs.push_back(std::move(q));
    }
    std::vector<void *> plans, outs;
    std::vector<const void *> ins;
    for (auto &q : qs) {
        plans.push_back(q.plan);
        ins.push_back(q.din);
        outs.push_back(q.dout);
    }
    if (syclfft_hip_execute_queue(plans.data(), ins.data(), outs.data(), qs.size(), nullptr, nullptr, err,
                                  sizeof(err)) != 0) {
        std::printf("FAIL queue execute: %s\n", err);
        return false;
    }
    bool all = true;
    for (auto &q : qs) {
        std::vector<double> got(2 * q.total), ref(2 * q.total);
        if (q.dbl) {
            std::memcpy(got.data(), q.dout, q.total * 16);
        } else {
            const float *f = static_cast<const float *>(q.dout);
            for (uint64_t k = 0; k < 2 * q.total; ++k) {
                got[k] = f[k];
            }
        }
        oracle_fftn_ld(static_cast<int>(q.lens.size()), q.lens.data(), q.batch, q.x.data(), ref.data(), q.dir,
                       q.norm);
        double num = 0, den = 0;
        for (uint64_t k = 0; k < 2 * q.total; ++k) {
            num += (got[k] - ref[k]) * (got[k] - ref[k]);
            den += ref[k] * ref[k];
        }
        const double rel = std::sqrt(num / (den > 0 ? den : 1));
        const double u_ = q.dbl ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
        const double tol = 4.0 * std::max(1.0, std::log2(static_cast<double>(q.total / q.batch))) * u_;
        const bool ok = rel <= tol && std::isfinite(rel);
        std::printf("%s %s relL2=%.3e tol=%.1e (queue)\n", ok ? "ok" : "FAIL", q.spec.c_str(), rel, tol);
        all = all && ok;
        syclfft_hip_plan_destroy(q.plan);
        hipFree(q.din);
        if (!q.inplace) {
            hipFree(q.dout);
        }
    }
    std::fflush(stdout);
    return all;
}

int main(int argc, char **argv)
{
    if (std::getenv("EMU_QUEUE")) {
        return run_queue(argc, argv) ? 0 : 1;
    }
    bool all = true;
    for (int i = 1; i < argc; ++i) {
        all = run_case(argv[i]) && all;
    }
    return all ? 0 : 1;
}
