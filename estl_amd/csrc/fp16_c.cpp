// fp16_c.cpp -- C ABI of the FP16-storage mode (kernels in fp16.hip).
//
// A plan created with ESTL_PRECISION_FP16 transforms buffers of binary16 (re, im) pairs
// that carry one FP32 scale per transform (true value = stored value * scale).
// estl_pack_fp16 / estl_unpack_fp16 convert from / to complex64 with per-transform
// max-abs scaling.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

static int log2_of(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

int execute_plan_fp16(Plan *p, const void *in, const float *in_scale, void *out, float *out_scale, hipStream_t st) {
    if (!in || !out || !in_scale || !out_scale) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (p->precision != ESTL_PRECISION_FP16) return set_error(ESTL_ERR_INVALID, "plan is not an FP16-storage plan");
    hipError_t e = hipSuccess;
    switch (p->kind) {
        case Plan::COPY:
            if (in != out) e = hipMemcpyAsync(out, in, 4 * (size_t)p->batch, hipMemcpyDeviceToDevice, st);
            break;
        case Plan::ROWS:
            e = launch_rows_fp16(p->log2n, p->dir, in, out, p->batch, p->tw_rows, p->scale, st);
            break;
        case Plan::PASSES: {
            const int P = (int)p->steps.size();
            float2 *ws0 = nullptr, *ws1 = nullptr;  // this stream's scratch
            int rc = stream_scratch(p, st, false, &ws0, nullptr);
            if (rc) return rc;
            std::vector<void *> dst(P);
            for (int q = P - 1, k = 0; q >= 0; --q, ++k) dst[q] = (k & 1) ? (void *)ws0 : out;
            if (P > 1 && dst[0] == in) {  // a single pass may run in place
                rc = stream_scratch(p, st, true, &ws0, &ws1);
                if (rc) return rc;
                dst[0] = ws1;
            }
            const void *src = in;
            for (int q = 0; q < P && e == hipSuccess; ++q) {
                e = launch_pass_fp16(p->steps[q].log2b, p->steps[q].log2t, p->dir, src, dst[q], p->steps[q].args, st);
                src = dst[q];
            }
            break;
        }
    }
    if (e == hipSuccess) e = launch_scale_mul(in_scale, out_scale, (float)p->out_factor, p->batch, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "FP16 transform launch: %s", hipGetErrorString(e));
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_pack_fp16(const void *d_x, void *d_x16, float *d_scale, long n, long batch, void *stream) {
    if (!d_x || !d_x16 || !d_scale) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (n < 1 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two", n);
    if (batch < 1) return set_error(ESTL_ERR_INVALID, "batch=%ld must be >= 1", batch);
    hipError_t e = launch_pack_fp16(static_cast<const float2 *>(d_x), d_x16, d_scale, log2_of(n), batch,
                                    static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "pack launch: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int estl_unpack_fp16(const void *d_x16, const float *d_scale, void *d_x, long n, long batch, void *stream) {
    if (!d_x || !d_x16 || !d_scale) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (n < 1 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two", n);
    if (batch < 1) return set_error(ESTL_ERR_INVALID, "batch=%ld must be >= 1", batch);
    hipError_t e = launch_unpack_fp16(d_x16, d_scale, static_cast<float2 *>(d_x), log2_of(n), batch,
                                      static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "unpack launch: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int estl_execute_fp16(estl_plan_t plan, const void *d_in, const float *d_in_scale, void *d_out, float *d_out_scale,
                      void *stream) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return execute_plan_fp16(reinterpret_cast<Plan *>(plan), d_in, d_in_scale, d_out, d_out_scale,
                             static_cast<hipStream_t>(stream));
}

}  // extern "C"
