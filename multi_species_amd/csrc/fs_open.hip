// Open-boundary field and the external drive fields.
//
// * E_MQ_DirectSum_openbcs::operator() (reference src/FieldStructure.cpp:
//   166-191): es[i] = sum_j q_ws[j] * 0.5 * z / sqrt(z^2 + eps^2) - x_i,
//   z = targets[i] - sources[j], no wrap (the -x_i is the neutralizing
//   background of unit density). FP64: a lane holds two targets, the sources
//   are the same for the whole wavefront (scalar loads), each run of 256
//   sources is summed in FP64 and the runs are added with TwoSum; small target
//   sets split the sources over workgroups and add the splits in a fixed
//   order, so the result does not depend on scheduling. When the targets are
//   the sources (one device array of 16384 points or more) the kernel's
//   oddness gives each pair once to both points (fs_sym.h, as the periodic
//   sum does): half the evaluations. Each pair costs a
//   subtraction, an FMA, an FP64 reciprocal square root (v_rsq_f64 plus a
//   third-order correction), a product and an FMA, all FP64 vector
//   instructions: the kernel is bound by FP64 VALU issue like the periodic one.
// * ExternalTanh / ExternalSine / ExternalLogistic::operator()
//   (FieldStructure.cpp:17-139): es[i] += a(t) sin(k x_i - omega t). The
//   amplitude a(t) is a scalar of the time alone, computed on the host with
//   the reference's own expressions and constants (FieldStructure.hpp:49-84);
//   the kernel adds the drive to every point.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fs_device.h"
#include "fs_internal.h"
#include "fs_sym.h"

namespace fs {
namespace {

constexpr int kBlock = 256;
constexpr int kTpt = 2;     // targets per lane
constexpr int kRun = 256;   // sources summed in plain FP64 before a TwoSum

// One (target block, source split) item: partial sums of the open-boundary
// kernel, 0.5 and the -x term applied by open_finish_kernel.
__global__ __launch_bounds__(kBlock) void open_f64_kernel(const double* __restrict__ tg, int64_t nt,
                                                          const double* __restrict__ src,
                                                          const double* __restrict__ w, int64_t ns,
                                                          int64_t chunk, double eps2,
                                                          double* __restrict__ part) {
    const int64_t i0 = (int64_t)blockIdx.x * kBlock * kTpt + threadIdx.x;
    const int64_t j_lo = (int64_t)blockIdx.y * chunk;
    const int64_t j_hi = std::min(ns, j_lo + chunk);
    double xi[kTpt], hi[kTpt], lo[kTpt];
#pragma unroll
    for (int t = 0; t < kTpt; ++t) {
        const int64_t i = i0 + (int64_t)t * kBlock;
        xi[t] = i < nt ? tg[i] : 0.0;
        hi[t] = 0.0;
        lo[t] = 0.0;
    }
    for (int64_t r0 = j_lo; r0 < j_hi; r0 += kRun) {
        const int64_t r1 = std::min(j_hi, r0 + kRun);
        double acc[kTpt];
#pragma unroll
        for (int t = 0; t < kTpt; ++t) acc[t] = 0.0;
        for (int64_t j = r0; j < r1; ++j) {
            const double xj = src[j], wj = w[j];
#pragma unroll
            for (int t = 0; t < kTpt; ++t) {
                const double z = xi[t] - xj;
                const double y = dev::rsq_f64(__builtin_fma(z, z, eps2));
                acc[t] = __builtin_fma(wj * z, y, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < kTpt; ++t) dev::two_sum_acc(hi[t], lo[t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < kTpt; ++t) {
        const int64_t i = i0 + (int64_t)t * kBlock;
        if (i < nt) part[(int64_t)blockIdx.y * nt + i] = hi[t] + lo[t];
    }
}

// es[i] = 0.5 * (sum of the splits, in split order) - x_i.
__global__ __launch_bounds__(kBlock) void open_finish_kernel(const double* __restrict__ part,
                                                             int64_t nt, int splits,
                                                             const double* __restrict__ tg,
                                                             double* __restrict__ es) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    double s = 0.0, lo = 0.0;
    for (int k = 0; k < splits; ++k) dev::two_sum_acc(s, lo, part[(int64_t)k * nt + i]);
    es[i] = 0.5 * (s + lo) - tg[i];
}

__global__ __launch_bounds__(kBlock) void external_kernel(double* __restrict__ es,
                                                          const double* __restrict__ tg,
                                                          int64_t nt, double a, double k,
                                                          double omt) {
    // the reference's rounding: k x and a sin(.) each rounded, no FMA
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    es[i] += a * sin(k * tg[i] - omt);
}

// Source splits: enough (target block, split) items for four workgroups per
// CU, each split at least 1024 sources. A function of the sizes and the
// device only.
int open_splits(int64_t nt, int64_t ns, int num_cus) {
    const int64_t blocks = (nt + kBlock * kTpt - 1) / (kBlock * kTpt);
    const int64_t want = (int64_t(4) * num_cus + blocks - 1) / std::max<int64_t>(blocks, 1);
    const int64_t most = std::max<int64_t>(1, ns / 1024);
    return (int)std::max<int64_t>(1, std::min<int64_t>({want, most, 65535}));
}

}  // namespace

// The drive amplitude of ExternalTanh (type 0, and any type other than 1
// and 2, as the reference deck loader's default branch), ExternalSine (1) or
// ExternalLogistic (2) at time t: FieldStructure.cpp:17-139 with the
// constants of FieldStructure.hpp:49-84.
double external_amplitude(int type, double t, double Am, double k) {
    if (type == 1) {  // ExternalSine: t1 = 50, t2 = 150, t3 = 200
        if (t < 50) return Am * std::sin(t * M_PI / 100);
        if (t < 150) return Am;
        if (t < 200) return Am * std::cos((t - 150) * M_PI / 100);
        return 0.0;
    }
    if (type == 2) {  // ExternalLogistic: t1 = 60
        if (t < 60) return Am / (1.0 + std::exp(-40.0 * (t - 10.0)));
        return Am * (1.0 - 1.0 / (1.0 + std::exp(-40.0 * (t - 110.0))));
    }
    // ExternalTanh: t0 = 0, tL = 69, tR = 307, twL = twR = 20; eps = drive(t0)
    auto drive = [](double s) { return 0.5 * (std::tanh((s - 69) / 20) - std::tanh((s - 307) / 20)); };
    const double eps = drive(0);
    return (drive(t) - eps) / (1 - eps) * Am * k;
}

int external_add(double* d_es, const double* d_targets, int64_t nt, double t, int type, double Am,
                 double k, double omega, hipStream_t stream) {
    if (nt <= 0) return FS_OK;
    const double a = external_amplitude(type, t, Am, k);
    external_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        d_es, d_targets, nt, a, k, omega * t);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

}  // namespace fs

extern "C" {

int fs_emq_open_f64_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                        const double* d_sources, const double* d_q_ws, int64_t ns,
                        double epsilon, void* stream) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (!(epsilon > 0)) return fs::fail(FS_E_INVALID, "epsilon must be > 0");
    if (nt == 0) return FS_OK;
    if (!d_es || !d_targets || (ns > 0 && (!d_sources || !d_q_ws)))
        return fs::fail(FS_E_INVALID, "null pointer");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (nt <= INT32_MAX && fs::use_sym(nt, ns, d_targets == d_sources, ctx->tuning)) {
        const fs::sym::Plan P = fs::sym::plan(nt, ctx->tuning, ctx->tuning.sym_budget, true);
        FS_TRY(ctx->scratch.ensure(fs::sym::scratch_bytes(P), st));
        FS_HIP_CHECK(fs::sym::run<true>(d_sources, d_q_ws, P, 0.0, 0.0, epsilon * epsilon,
                                        ctx->scratch.ptr, d_es, st));
        ctx->f64_pair_evals = fs::sym::evals(P);
        return FS_OK;
    }
    ctx->f64_pair_evals = nt * ns;
    const int splits = fs::open_splits(nt, ns, ctx->num_cus);
    const int64_t chunk = ns > 0 ? (ns + splits - 1) / splits : 0;
    FS_TRY(ctx->scratch.ensure(sizeof(double) * (size_t)splits * (size_t)nt, st));
    double* part = ctx->scratch.as<double>(0);
    const int64_t blocks = (nt + fs::kBlock * fs::kTpt - 1) / (fs::kBlock * fs::kTpt);
    if (blocks > 0x7fffffff) return fs::fail(FS_E_INVALID, "too many targets");
    fs::open_f64_kernel<<<dim3((unsigned)blocks, (unsigned)splits), fs::kBlock, 0, st>>>(
        d_targets, nt, d_sources, d_q_ws, ns, chunk, epsilon * epsilon, part);
    FS_HIP_CHECK(hipGetLastError());
    fs::open_finish_kernel<<<(unsigned)((nt + fs::kBlock - 1) / fs::kBlock), fs::kBlock, 0, st>>>(
        part, nt, splits, d_targets, d_es);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_emq_open_f64(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                    const double* sources, const double* q_ws, int64_t ns, double epsilon) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (!(epsilon > 0)) return fs::fail(FS_E_INVALID, "epsilon must be > 0");
    if (nt == 0) return FS_OK;
    if (!es || !targets || (ns > 0 && (!sources || !q_ws)))
        return fs::fail(FS_E_INVALID, "null pointer");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bt = fs::align_up(sizeof(double) * nt, 256);
    const size_t bs = fs::align_up(sizeof(double) * ns, 256);
    FS_TRY(ctx->io.ensure(2 * bt + 2 * bs, ctx->stream));
    double* d_t = ctx->io.as<double>(0);
    double* d_e = ctx->io.as<double>(bt);
    double* d_s = ctx->io.as<double>(2 * bt);
    double* d_w = ctx->io.as<double>(2 * bt + bs);
    // targets and sources the same positions (one array, or a copy of it):
    // one device copy, so the sum takes each pair once
    const bool same = nt == ns && (targets == sources ||
                                   std::memcmp(targets, sources, sizeof(double) * nt) == 0);
    if (same)
        d_t = d_s;
    else
        FS_HIP_CHECK(hipMemcpyAsync(d_t, targets, sizeof(double) * nt, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (ns > 0) {
        FS_HIP_CHECK(hipMemcpyAsync(d_s, sources, sizeof(double) * ns, hipMemcpyHostToDevice,
                                    ctx->stream));
        FS_HIP_CHECK(hipMemcpyAsync(d_w, q_ws, sizeof(double) * ns, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    FS_TRY(fs_emq_open_f64_dev(ctx, d_e, d_t, nt, d_s, d_w, ns, epsilon, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

int fs_external_field_add_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                              double t, int type, double Am, double k, double omega,
                              void* stream) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (nt > 0 && (!d_es || !d_targets)) return fs::fail(FS_E_INVALID, "null pointer");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return fs::external_add(d_es, d_targets, nt, t, type, Am, k, omega, st);
}

int fs_external_field_add(fs_ctx* ctx, double* es, const double* targets, int64_t nt, double t,
                          int type, double Am, double k, double omega) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (nt == 0) return FS_OK;
    if (!es || !targets) return fs::fail(FS_E_INVALID, "null pointer");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bt = fs::align_up(sizeof(double) * nt, 256);
    FS_TRY(ctx->io.ensure(2 * bt, ctx->stream));
    double* d_t = ctx->io.as<double>(0);
    double* d_e = ctx->io.as<double>(bt);
    FS_HIP_CHECK(hipMemcpyAsync(d_t, targets, sizeof(double) * nt, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_e, es, sizeof(double) * nt, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_TRY(fs::external_add(d_e, d_t, nt, t, type, Am, k, omega, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

}  // extern "C"
