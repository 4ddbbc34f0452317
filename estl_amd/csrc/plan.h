// plan.h -- host-side plan object behind estl_plan_t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"

namespace estl {

struct PassStep {
    int log2b;   // radix 2^log2b
    int log2t;   // 2^log2t adjacent columns per workgroup
    PassArgs args;
};

struct Plan {
    enum Kind { COPY, ROWS, PASSES };
    int device = 0;
    long n = 0, batch = 0;
    int dir = -1, precision = 0, log2n = 0;
    unsigned flags = 0;
    float scale = 1.0f;
    double out_factor = 1.0;  // FP16 storage: output scale = input scale * out_factor
    Kind kind = COPY;
    const float2 *tw_rows = nullptr;
    std::vector<PassStep> steps;
    size_t scratch_bytes = 0;
    float2 *scratch[2] = {nullptr, nullptr};  // the plan's own (profiling, synchronous paths)
    // scratch of executes without caller workspace, one pair per stream, so executes of
    // one plan on different streams never share a buffer
    struct StreamScratch {
        std::mutex mu;
        std::map<hipStream_t, std::pair<float2 *, float2 *>> bufs;
    };
    std::shared_ptr<StreamScratch> by_stream = std::make_shared<StreamScratch>();
    // resident buffers and stream of the synchronous host entry point
    float2 *h_in_dev = nullptr, *h_out_dev = nullptr;
    hipStream_t stream = nullptr;
    // mean time of each launch from the last estl_plan_profile
    std::vector<double> kernel_ms;
};

int set_error(int code, const char *fmt, ...);

// bytes that `count` complex values span in the tile-planarized layout: up to the
// last value's imaginary part (whole tiles when count is a multiple of 32; a partial
// last tile ends early, so a host array sized by the reference's
// PlannarizedComplexArray rounding is never overrun)
inline size_t planar_bytes(long count) {
    if (count < 1) return 0;
    const long last = count - 1;
    return (size_t)((((last >> 5) << 6) | (last & 31)) + 33) * sizeof(float);
}

// return ESTL_ERR_HIP with the failing call's text from the enclosing function
#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return set_error(ESTL_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                             __FILE__, __LINE__);                                                  \
    } while (0)

int current_device();
int block_table(int device, int log2n, const float2 **out);
int power_table(int device, int log2n, const float2 **out);  // block_fft's kTwPowers table
int linear_table(int device, long count, long mult, long nq, const float2 **out);
std::vector<int> default_pass_split(int log2n);
int build_plan(Plan *p, long n, long batch, int dir, int precision, unsigned flags, const std::vector<int> *split,
               const std::vector<int> *tiles = nullptr);
int execute_plan(Plan *p, const float2 *in, float2 *out, hipStream_t st);
int execute_plan_ws(Plan *p, const float2 *in, float2 *out, float2 *ws0, float2 *ws1, hipStream_t st);
int ensure_scratch(Plan *p, int which);
// scratch pair of stream st (allocated on first use; second only if `two`)
int stream_scratch(Plan *p, hipStream_t st, bool two, float2 **ws0, float2 **ws1);
void destroy_plan(Plan *p);
// device address of the no-allocation entry point's static scratch on the current device
// (noalloc.hip; looked up once per device)
int noalloc_scratch(float2 **out);
void free_stream_scratch(Plan *p);  // the per-stream scratch of a plan not destroyed by destroy_plan
std::string describe_plan(const Plan *p);
int plan_launches(const Plan *p);
int tune_plan(long n, long batch, int dir, int precision, bool *rows, std::vector<int> *split, std::vector<int> *tiles,
              std::string *report);

// ---- batched circular convolution: immutable shape + per-execution buffers
struct ConvShape {
    int precision = 0;  // ESTL_PRECISION_FP32 or _FP16 (binary16 rows, per-row scales)
    int log2n = 0;
    long n = 0, rows = 0, pairs = 0;
    unsigned flags = 0;
    bool corr = false, fused = false;
    float scale = 1.0f;
    const float2 *tw_rows = nullptr;  // power table of the fused rows kernel (n <= 2^13)
    Plan spec_plan;                  // forward FFT of the folded filter (batch 1)
    const float2 *tw_spec = nullptr;  // k_filter_spectrum's table when spec_plan is a rows plan
    std::vector<PassStep> fwd, inv;  // multi-pass pipeline (n > 2^13): fwd, mid, inv
    PassStep mid{};                  // last forward + multiply + first inverse pass
    size_t scratch_bytes = 0;        // bytes of each of the two scratch buffers
    size_t spec_bytes = 0;           // bytes of the spectrum (FP16: n + 1 values, the
                                     // last holding the filter's L1 bound)
    // one FP32 row of n >= 2^14 (conv_half_path): transforms of n/2 values around
    // k_conv_half_merge, taken when the row's input and output are 8-byte aligned
    bool half = false;
    Plan half_fwd, half_inv;
    // tables of the overlap-save windows of 2^13 (n > 2^13) and 2^14 (n > 2^14) values:
    // block tables (the filter's window spectrum, kSpecTw) and power tables (k_conv_ols)
    const float2 *tw_ols[2] = {nullptr, nullptr};
    const float2 *tw_ols_pow[2] = {nullptr, nullptr};
};
bool conv_half_path(long n, long rows, int precision);
// Rows longer than an overlap-save window convolved with a filter of at most
// kOlsWideMaxTaps taps run overlap-save unless ESTL_CONV_FULL_LENGTH asks for whole-row
// transforms: one launch (conv.hip k_conv_ols) over windows of L = 2^13 values (K <= 4097
// taps; 512-thread workgroups, two per CU) or 2^14 values (n >= 2^15, K <= 8193; 1024
// threads, one per CU), each giving M = L - K + 1 outputs, with the filter's L-point
// spectrum.  Returns log2 L, or 0 for whole-row transforms.
//
// Between the two window lengths (both possible) the choice is a cost model fitted to
// MI355X timings of the two kernels (68 shapes of 2^22 to 2^28 values, K = 16..4097; the
// choice is within 1.5 % of the faster length at each, where the two differ by up to
// 30 %).  Window pairs are dealt to the `cus` CUs and the time is that of the CU with the
// most: with T the time a CU takes for two concurrent 2^13 pairs (four windows), a lone
// 2^13 pair takes 0.35 T and a 2^14 pair (one per CU) 1.10 T -- or 0.90 T when every CU
// has at most one (a single short round).  cus = 0: the windows' count and size alone.
// 2^13 windows read fewer padding values for short filters (L/M near 1 for either length,
// and the last window of a row less wasted); 2^14 windows read each input value fewer
// times for long filters (L/M = 1.14 against 1.33 at K = 2048).
inline int conv_ols_log2(long n, long rows, long k, unsigned flags, int cus) {
    if (k < 1 || rows < 1 || (flags & ESTL_CONV_FULL_LENGTH)) return 0;
    const bool narrow = n > (1L << kOlsLog2) && k <= kOlsMaxTaps;
    const bool wide = n > (1L << kOlsWideLog2) && k <= kOlsWideMaxTaps;
    if (!narrow || !wide) return narrow ? kOlsLog2 : wide ? kOlsWideLog2 : 0;
    const long o13 = ols_outputs(kOlsLog2, k, false), o14 = ols_outputs(kOlsWideLog2, k, false);  // per window
    const long w13 = rows * ((n + o13 - 1) / o13), w14 = rows * ((n + o14 - 1) / o14);          // windows
    double t13, t14;
    if (cus > 0) {
        const long m13 = ((w13 + 1) / 2 + cus - 1) / cus, m14 = ((w14 + 1) / 2 + cus - 1) / cus;
        t13 = (double)(m13 / 2) + 0.35 * (double)(m13 % 2);
        t14 = (m14 == 1 ? 0.90 : 1.10) * (double)m14;
    } else {
        t13 = (double)w13 / 4.0;
        t14 = 1.10 * (double)w14 / 2.0;
    }
    return t14 < t13 ? kOlsWideLog2 : kOlsLog2;
}
// Both window lengths possible (the choice then depends on the rows of the call): a
// filter registered once keeps both spectra, the 2^14 one at conv_ols_wide_offset
// (float2 values) after the 2^13 one (FP16 storage: each followed by its L1 pair).
inline bool conv_ols_both(long n, long k, unsigned flags) {
    return !(flags & ESTL_CONV_FULL_LENGTH) && n > (1L << kOlsWideLog2) && k >= 1 && k <= kOlsMaxTaps;
}
inline long conv_ols_wide_offset(bool fp16) { return (1L << kOlsLog2) + (fp16 ? 1 : 0); }
// compute units of the current device (cached per device; conv_c.cpp)
int device_cus();
int conv_run_group_half(const Plan &fwd, const Plan &inv, const ConvGroup *d_group, int members, long n, bool corr,
                        float2 *z, float2 *ws, hipStream_t st);
struct ConvBuffers {
    float2 *spec = nullptr, *dc = nullptr;
    float2 *scratch[2] = {nullptr, nullptr};
    long k = 0;  // taps of the filter whose spectrum `spec` holds (picks the pipeline)
    int ols = 0;  // log2 of the overlap-save window of that spectrum (conv_ols_log2), 0: whole rows
};
int conv_shape_init(ConvShape *c, long n, long rows, unsigned flags, int precision = 0);
size_t conv_footprint(const ConvShape &c);  // bytes of ConvBuffers
// bytes of each of the two scratch buffers: the pipeline's, or more if the filter's
// (FP32) spectrum transform, which borrows scratch[0], needs more
size_t conv_scratch_bytes(const ConvShape &c);
int conv_filter(const ConvShape &c, const ConvBuffers &b, const float *h, long k, hipStream_t st);
// request groups (the server's continuous batching): the spectrum plan over `filters`
// rows, the group's spectra and its convolution passes
int group_spectrum_plan(Plan *p, long n, long filters);
// skip_first: every member's filter passes spectrum_skips_first_pass (its k, its shift)
bool spectrum_skips_first_pass(const Plan &spec, long k, long shift);
// divide_l1 (FP16 storage): the spectra stored divided by the members' L, which
// conv_fp16_filter_group wrote into the descriptor beforehand
int conv_filter_group(const Plan &spec, const ConvGroup *d_group, float2 *spectra, float2 *scratch0,
                      bool skip_first, bool divide_l1, hipStream_t st);
int conv_filter_group_rows(const ConvShape &c, const ConvGroup *d_group, int members, bool divide_l1,
                           hipStream_t st);
int conv_fp16_filter_group(ConvGroup *d_group, int members, hipStream_t st);
// registered: the members' spectra are registered filters' (already divided by L, which
// follows their n values)
int conv_run_group_fp16(const ConvShape &c, ConvGroup *d_group, int members, long pairs, long max_rows,
                        bool registered, void *scratch0, void *scratch1, hipStream_t st);
int conv_run_group(const ConvShape &c, const ConvGroup *d_group, long pairs, float2 *scratch0, float2 *scratch1,
                   hipStream_t st);
int conv_run(const ConvShape &c, const ConvBuffers &b, const float *x, float *y, hipStream_t st);
// FP16 storage: x16 / y16 binary16 rows, y = (stored y) * s_out[row]
int conv_run_fp16(const ConvShape &c, const ConvBuffers &b, const void *x16, const float *s_in, void *y16,
                  float *s_out, hipStream_t st);

}  // namespace estl
