// noalloc.hip -- estl_execute_device_noalloc: estl_execute_device without any device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// estl_execute_device takes a multi-pass plan's scratch from a per-stream pair that it
// allocates on a stream's first use.  This variant takes it from `scratch`, a static
// device array of kNoAllocScratchBytes per GPU (held until the process ends), whose
// device address is fetched once per device with hipGetSymbolAddress (at estl_init, or
// at the first call on a device) and used from then on.  Every pass writes all of its
// output before the next pass reads it, so nothing in the array has to start at zero.
// All calls on one GPU share its array: calls that overlap in time on one GPU (other
// streams, other threads) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/estl_hip.h"
#include "plan.h"

namespace estl {

// 64 MiB: 2^23 complex64 values -- one scratch buffer for plans of n * batch <= 2^23
// values, or two (an in-place chain of an odd number of passes) of 2^22 each
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

static __device__ __align__(256) float2 scratch[kNoAllocScratchBytes / sizeof(float2)];

namespace {
std::mutex g_mu;
constexpr int kMaxDevices = 64;
float2 *g_ptr[kMaxDevices] = {};
}  // namespace

// The array's device address on the current device (looked up once per device).
int noalloc_scratch(float2 **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(ESTL_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(ESTL_ERR_HIP, "no device address for the no-allocation scratch");
        g_ptr[dev] = static_cast<float2 *>(p);
    }
    *out = g_ptr[dev];
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" int estl_execute_device_noalloc(estl_plan_t plan, const void *d_in, void *d_out, void *stream) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    Plan *p = reinterpret_cast<Plan *>(plan);
    if (!d_in || !d_out) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (p->precision != ESTL_PRECISION_FP32)
        return set_error(ESTL_ERR_INVALID, "estl_execute_device_noalloc runs FP32 plans");
    float2 *ws0 = nullptr, *ws1 = nullptr;
    if (p->kind == Plan::PASSES) {
        // an in-place chain of an odd number of passes needs a second buffer
        const bool two = p->steps.size() > 1 && (p->steps.size() & 1) && d_in == d_out;
        const size_t need = p->scratch_bytes * (two ? 2 : 1);
        if (need > kNoAllocScratchBytes)
            return set_error(ESTL_ERR_INVALID,
                             "plan of %ld x %ld values needs %zu bytes of scratch; the no-allocation entry point "
                             "holds %zu",
                             p->n, p->batch, need, kNoAllocScratchBytes);
        float2 *base = nullptr;
        if (int rc = noalloc_scratch(&base)) return rc;
        ws0 = base;
        if (two) ws1 = base + p->scratch_bytes / sizeof(float2);
    }
    return execute_plan_ws(p, static_cast<const float2 *>(d_in), static_cast<float2 *>(d_out), ws0, ws1,
                           static_cast<hipStream_t>(stream));
}
