// profile.cpp -- per-launch timing of a plan: the surface the reference's harness reads
// from an OpenCL backend (kernelTimers, getTotalComputationFlops(q), getKernelInfo(q):
// FFT_OpenCL_Contiguous.h:73-78,94-95, uOpenCL.h:281-282,339-340) and prints as
// "[K<q> (<info>): <GFLOP/s> Gfs (<share>%)]" (tests/Test.cpp:86-100).
//
// Each launch of the plan runs on its own between two HIP events on a private stream,
// `reps` times after one warm-up; the mean per launch is kept in the plan.  Flops
// follow the reference's convention: 5 n log2(B_q) per transform for a pass of radix
// B_q (the rows kernel: 5 n log2 n).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

static int ilog2l(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

int plan_kernel_count(const Plan *p) {
    switch (p->kind) {
        case Plan::COPY: return 0;
        case Plan::ROWS: return 1;
        case Plan::PASSES: return (int)p->steps.size();
    }
    return 0;
}

double plan_kernel_flops(const Plan *p, int q) {
    if (q < 0 || q >= plan_kernel_count(p)) return -1.0;
    const int bits = p->kind == Plan::ROWS ? p->log2n : p->steps[q].log2b;
    return 5.0 * (double)p->n * bits * (double)p->batch;
}

std::string plan_kernel_info(const Plan *p, int q) {
    char buf[128];
    if (q < 0 || q >= plan_kernel_count(p)) return "";
    const bool h = p->precision == ESTL_PRECISION_FP16;
    if (p->kind == Plan::ROWS) {
        std::string r;
        for (int k = 0; k < rows_pass_count(p->log2n); ++k) {
            long A;
            int B;
            rows_pass_params(p->log2n, k, &A, &B);
            snprintf(buf, sizeof buf, "%s%d", k ? "x" : "", B);
            r += buf;
        }
        snprintf(buf, sizeof buf, "%s, radices %s in registers", h ? "k_fft_rows_h" : "k_fft_rows", r.c_str());
        return buf;
    }
    const PassStep &s = p->steps[q];
    snprintf(buf, sizeof buf, "%s, radix %ld, %d columns per workgroup", h ? "k_fft_pass_h" : "k_fft_pass",
             1L << s.log2b, 1 << s.log2t);
    return buf;
}

// One launch q of the plan from src to dst (the same launches execute_plan makes).
static hipError_t launch_one(Plan *p, int q, const void *src, void *dst, hipStream_t st) {
    const bool h = p->precision == ESTL_PRECISION_FP16;
    if (p->kind == Plan::ROWS) {
        if (h) return launch_rows_fp16(p->log2n, p->dir, src, dst, p->batch, p->tw_rows, p->scale, st);
        auto f = (p->flags & ESTL_FLAG_PLANAR) ? launch_rows_planar : launch_rows;
        return f(p->log2n, p->dir, static_cast<const float2 *>(src), static_cast<float2 *>(dst), p->batch, p->tw_rows,
                 p->scale, st);
    }
    const PassStep &s = p->steps[q];
    return h ? launch_pass_fp16(s.log2b, s.log2t, p->dir, src, dst, s.args, st)
             : launch_pass(s.log2b, s.log2t, p->dir, static_cast<const float2 *>(src), static_cast<float2 *>(dst),
                           s.args, st);
}

int profile_plan(Plan *p, const void *in, void *out, int reps, std::string *json) {
    const int K = plan_kernel_count(p);
    if (reps < 1) reps = 1;
    const size_t elem = p->precision == ESTL_PRECISION_FP16 ? 4 : sizeof(float2);
    const size_t bytes = (p->flags & ESTL_FLAG_PLANAR) ? planar_bytes(p->n * p->batch) : elem * (size_t)p->n * (size_t)p->batch;
    if (!p->stream) HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    if (!in || !out) {
        // the same sizes estl_execute_host gives these buffers
        const size_t hb = (p->flags & ESTL_FLAG_PLANAR) ? planar_bytes(p->n * p->batch)
                                                        : sizeof(float2) * (size_t)p->n * (size_t)p->batch;
        if (!p->h_in_dev) HIP_TRY(hipMalloc(&p->h_in_dev, hb));
        if (!p->h_out_dev) HIP_TRY(hipMalloc(&p->h_out_dev, hb));
        HIP_TRY(hipMemsetAsync(p->h_in_dev, 0, bytes, p->stream));
        in = p->h_in_dev;
        out = p->h_out_dev;
    }
    if (p->kind == Plan::PASSES) {
        int rc = ensure_scratch(p, 0);
        if (!rc) rc = ensure_scratch(p, 1);
        if (rc) return rc;
    }
    // source / destination of each launch: in -> s0 -> s1 -> ... -> out (never in place)
    std::vector<const void *> src(K);
    std::vector<void *> dst(K);
    for (int q = 0; q < K; ++q) {
        src[q] = q == 0 ? in : (const void *)p->scratch[(q - 1) & 1];
        dst[q] = q == K - 1 ? out : (void *)p->scratch[q & 1];
    }
    std::vector<hipEvent_t> ev(K + 1);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    std::vector<double> ms(K, 0.0);
    hipError_t err = hipSuccess;
    for (int r = 0; r <= reps && err == hipSuccess; ++r) {
        HIP_TRY(hipEventRecord(ev[0], p->stream));
        for (int q = 0; q < K && err == hipSuccess; ++q) {
            err = launch_one(p, q, src[q], dst[q], p->stream);
            if (err == hipSuccess) err = hipEventRecord(ev[q + 1], p->stream);
        }
        if (err != hipSuccess) break;
        HIP_TRY(hipEventSynchronize(ev[K]));
        if (r == 0) continue;  // warm-up
        for (int q = 0; q < K; ++q) {
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
            ms[q] += t / reps;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (err != hipSuccess) return set_error(ESTL_ERR_HIP, "profile launch: %s", hipGetErrorString(err));
    p->kernel_ms = ms;
    double total = 0, flops = 0;
    for (int q = 0; q < K; ++q) {
        total += ms[q];
        flops += plan_kernel_flops(p, q);
    }
    char buf[512];
    std::string s = "{\"kernels\": [";
    for (int q = 0; q < K; ++q) {
        const double f = plan_kernel_flops(p, q);
        snprintf(buf, sizeof buf,
                 "%s{\"index\": %d, \"info\": \"%s\", \"ms\": %.6f, \"flops\": %.0f, \"gflops\": %.3f, \"share\": %.9f}",
                 q ? ", " : "", q, plan_kernel_info(p, q).c_str(), ms[q], f, ms[q] > 0 ? f / (ms[q] * 1e6) : 0.0,
                 total > 0 ? ms[q] / total : 0.0);
        s += buf;
    }
    snprintf(buf, sizeof buf, "], \"total_ms\": %.6f, \"total_flops\": %.0f, \"total_gflops\": %.3f, \"reps\": %d}", total,
             flops, total > 0 ? flops / (total * 1e6) : 0.0, reps);
    s += buf;
    *json = s;
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_plan_profile(estl_plan_t plan, const void *d_in, void *d_out, int reps, char *json, int len) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    std::string s;
    int rc = profile_plan(reinterpret_cast<Plan *>(plan), d_in, d_out, reps, &s);
    if (rc) return rc;
    if (json && len > 0) {
        strncpy(json, s.c_str(), (size_t)len - 1);
        json[len - 1] = 0;
    }
    return (int)s.size();
}

double estl_plan_kernel_flops(estl_plan_t plan, int q) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return plan_kernel_flops(reinterpret_cast<Plan *>(plan), q);
}

double estl_plan_kernel_ms(estl_plan_t plan, int q) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    const Plan *p = reinterpret_cast<Plan *>(plan);
    if (q < 0 || q >= (int)p->kernel_ms.size()) return -1.0;
    return p->kernel_ms[q];
}

int estl_plan_kernel_info(estl_plan_t plan, int q, char *buf, int len) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    const std::string s = plan_kernel_info(reinterpret_cast<Plan *>(plan), q);
    if (s.empty()) return set_error(ESTL_ERR_INVALID, "kernel index %d out of range", q);
    if (buf && len > 0) {
        strncpy(buf, s.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)s.size();
}

}  // extern "C"
