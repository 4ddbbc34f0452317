// conv_c.cpp -- C ABI of the batched 1-D circular convolution (Convolution.h:95-148
// generalised: real rows of length n, a real filter of any length, batch of rows).
//
// Plan: n <= 2^13 -> k_conv_rows (one launch: pair-pack, forward FFT, spectrum
// multiply, inverse FFT, scale, unpack); larger n -> the forward global passes with the
// first reading real pairs and the last multiplying by the spectrum, then the inverse
// passes with the last scaling and unpacking (pass_tile.h I/O modes).  The filter
// spectrum is computed once per estl_conv_set_filter and stays resident.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

hipError_t launch_conv_rows(int log2n, bool corr, const float *x, float *y, long rows, const float2 *spec,
                            const float2 *tw, float scale, hipStream_t st);
hipError_t launch_conv_fold(const float *h, long k, long n, long shift, float2 *hp, hipStream_t st);
hipError_t launch_spec_normalize(float2 *spec, long n, const float2 *dc, hipStream_t st);

struct Conv {
    int device = 0, log2n = 0;
    long n = 0, rows = 0, pairs = 0;
    unsigned flags = 0;
    bool corr = false, fused_rows = false, have_filter = false;
    float scale = 1.0f;
    const float2 *tw_rows = nullptr;
    float2 *spec = nullptr, *hp = nullptr, *dc = nullptr;
    Plan spec_plan;                  // forward FFT of the folded filter (batch 1)
    std::vector<PassStep> fwd, inv;  // multi-pass pipeline
    float2 *scratch[2] = {nullptr, nullptr};
    size_t scratch_bytes = 0;
};

static void conv_free(Conv *c) {
    if (!c) return;
    if (c->spec) (void)hipFree(c->spec);
    if (c->hp) (void)hipFree(c->hp);
    if (c->dc) (void)hipFree(c->dc);
    for (auto &s : c->scratch)
        if (s) (void)hipFree(s);
    for (auto &s : c->spec_plan.scratch)
        if (s) (void)hipFree(s);
    delete c;
}

static int conv_create(Conv **out, long n, long rows, int precision, unsigned flags) {
    *out = nullptr;
    if (n < 2 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two >= 2", n);
    if (rows < 1) return set_error(ESTL_ERR_INVALID, "rows=%ld must be >= 1", rows);
    if (precision != ESTL_PRECISION_FP32)
        return set_error(ESTL_ERR_UNSUPPORTED, "convolution precision %d not supported", precision);
    const unsigned known = ESTL_CONV_CORRELATE | ESTL_CONV_NORMALIZE_KERNEL | ESTL_CONV_CENTER;
    if (flags & ~known) return set_error(ESTL_ERR_INVALID, "unknown convolution flags 0x%x", flags);
    Conv *c = new Conv();
    c->device = current_device();
    c->n = n;
    c->rows = rows;
    c->pairs = (rows + 1) / 2;
    c->flags = flags;
    c->corr = (flags & ESTL_CONV_CORRELATE) != 0;
    while ((1L << c->log2n) < n) ++c->log2n;
    c->scale = 1.0f / (float)n;
    int rc = build_plan(&c->spec_plan, n, 1, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
    if (rc) {
        conv_free(c);
        return rc;
    }
    if (hipMalloc(&c->spec, sizeof(float2) * n) != hipSuccess || hipMalloc(&c->hp, sizeof(float2) * n) != hipSuccess ||
        hipMalloc(&c->dc, sizeof(float2)) != hipSuccess) {
        conv_free(c);
        return set_error(ESTL_ERR_NOMEM, "convolution: spectrum buffers (%ld values)", n);
    }
    if (c->log2n <= kMaxRowsLog2) {
        c->fused_rows = true;
        rc = block_table(c->device, c->log2n, &c->tw_rows);
        if (rc) {
            conv_free(c);
            return rc;
        }
    } else {
        Plan f, b;
        rc = build_plan(&f, n, c->pairs, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
        if (!rc) rc = build_plan(&b, n, c->pairs, ESTL_BACKWARD, ESTL_PRECISION_FP32, 0, nullptr);
        if (rc) {
            conv_free(c);
            return rc;
        }
        c->fwd = f.steps;
        c->inv = b.steps;
        // the first pass reads real pairs through its direct (S >= T) path
        PassStep &f0 = c->fwd.front();
        while ((1 << f0.log2t) > (1L << f0.args.log2S) && f0.log2t > 0) --f0.log2t;
        f0.args.io_in = 1;
        f0.args.real_rows = rows;
        c->fwd.back().args.io_out = 1;
        c->fwd.back().args.corr = c->corr ? 1 : 0;
        c->inv.back().args.io_out = 2;
        c->inv.back().args.real_rows = rows;
        c->scratch_bytes = sizeof(float2) * (size_t)n * (size_t)c->pairs;
        for (auto &s : c->scratch)
            if (hipMalloc(&s, c->scratch_bytes) != hipSuccess) {
                conv_free(c);
                return set_error(ESTL_ERR_NOMEM, "convolution scratch (%zu bytes)", c->scratch_bytes);
            }
    }
    *out = c;
    return ESTL_OK;
}

static int conv_set_filter(Conv *c, const float *h, long k, hipStream_t st) {
    if (!h || k < 1) return set_error(ESTL_ERR_INVALID, "filter needs k >= 1 values");
    const long shift = (c->flags & ESTL_CONV_CENTER) ? ((k + 1) / 2) % c->n : 0;
    hipError_t e = launch_conv_fold(h, k, c->n, shift, c->hp, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "fold launch: %s", hipGetErrorString(e));
    int rc = execute_plan(&c->spec_plan, c->hp, c->spec, st);
    if (rc) return rc;
    if (c->flags & ESTL_CONV_NORMALIZE_KERNEL) {
        // divide by the kernel sum K^[0] (Convolution.h:141): spectrum / spectrum[0].re
        e = hipMemcpyAsync(c->dc, c->spec, sizeof(float2), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "spectrum dc copy: %s", hipGetErrorString(e));
        e = launch_spec_normalize(c->spec, c->n, c->dc, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "normalize launch: %s", hipGetErrorString(e));
    }
    c->have_filter = true;
    return ESTL_OK;
}

static int conv_execute(Conv *c, const float *x, float *y, hipStream_t st) {
    if (!x || !y) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (!c->have_filter) return set_error(ESTL_ERR_INVALID, "estl_conv_set_filter has not been called");
    if (c->fused_rows) {
        hipError_t e = launch_conv_rows(c->log2n, c->corr, x, y, c->rows, c->spec, c->tw_rows, c->scale, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "k_conv_rows launch: %s", hipGetErrorString(e));
        return ESTL_OK;
    }
    // forward passes: x (real pairs) -> s0 -> s1 -> ..., the last multiplies by the
    // spectrum; inverse passes continue the ping-pong, the last writes y (real pairs)
    int cur = 0;
    const float2 *src = c->scratch[1];  // not read by the first pass (it reads x)
    auto run = [&](PassStep st_, int dir) -> int {
        PassArgs a = st_.args;
        a.rin = x;
        a.rout = y;
        a.spectrum = c->spec;
        if (a.io_out == 2) a.scale = c->scale;
        float2 *dst = c->scratch[cur];
        hipError_t e = launch_pass(st_.log2b, st_.log2t, dir, src, dst, a, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "conv pass launch: %s", hipGetErrorString(e));
        src = dst;
        cur ^= 1;
        return ESTL_OK;
    };
    for (auto &s : c->fwd)
        if (int rc = run(s, ESTL_FORWARD)) return rc;
    for (auto &s : c->inv)
        if (int rc = run(s, ESTL_BACKWARD)) return rc;
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_conv_create(estl_conv_t *conv, long n, long rows, int precision, unsigned flags) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle pointer is null");
    *conv = nullptr;
    Conv *c = nullptr;
    int rc = conv_create(&c, n, rows, precision, flags);
    if (rc) return rc;
    *conv = reinterpret_cast<estl_conv_t>(c);
    return ESTL_OK;
}

int estl_conv_set_filter(estl_conv_t conv, const float *d_h, long k, void *stream) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return conv_set_filter(reinterpret_cast<Conv *>(conv), d_h, k, static_cast<hipStream_t>(stream));
}

int estl_conv_execute(estl_conv_t conv, const float *d_x, float *d_y, void *stream) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return conv_execute(reinterpret_cast<Conv *>(conv), d_x, d_y, static_cast<hipStream_t>(stream));
}

long long estl_conv_workspace_bytes(estl_conv_t conv) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    Conv *c = reinterpret_cast<Conv *>(conv);
    return (long long)(2 * c->scratch_bytes + 3 * sizeof(float2) * (size_t)c->n);
}

int estl_conv_destroy(estl_conv_t conv) {
    conv_free(reinterpret_cast<Conv *>(conv));
    return ESTL_OK;
}

}  // extern "C"
