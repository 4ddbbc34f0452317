// conv_c.cpp -- batched 1-D circular convolution (Convolution.h:95-148 generalised:
// real rows of length n, a real filter of any length, a batch of rows) and its C ABI.
//
// A ConvShape (immutable per n, rows, flags) picks the pipeline:
//   n <= 2^13: k_conv_rows, one launch (pair-pack, forward FFT, spectrum multiply,
//              inverse FFT, scale, unpack);
//   n >  2^13: the forward global passes (the first reading real pairs), one middle
//              pass fusing the last forward pass, the spectrum multiply and the first
//              inverse pass, then the remaining inverse passes (the last scaling and
//              unpacking): 2P - 1 launches for a P-pass split (pass_tile.h I/O modes);
//   one FP32 row, n >= 2^14: the row read as n/2 complex values (even and odd samples),
//              a forward plan of n/2, k_conv_half_merge (the row's spectrum from the
//              half-length one, the multiply, and back), an inverse plan of n/2 --
//              half the values of a pair whose second row is empty.
// ConvBuffers (spectrum, folded filter, two scratch buffers) are separate so a caller
// can place them: the estl_conv_* object owns its own, the server takes them from its
// HBM arena.  The filter spectrum is computed once per filter and stays resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

// (group: a request group's descriptor in device memory, or null)
hipError_t launch_conv_rows(int log2n, bool corr, const float *x, float *y, long rows, const float2 *spec,
                            const float2 *tw, float scale, hipStream_t st, const ConvGroup *group = nullptr);
hipError_t launch_filter_spectrum(int log2n, const float *h, long k, long shift, float2 *spec, const float2 *tw,
                                  bool normalize, hipStream_t st, const ConvGroup *group = nullptr, int members = 0,
                                  bool divide_l1 = false, float2 *dc = nullptr, float2 *l1_single = nullptr);
hipError_t launch_spec_normalize(float2 *spec, long n, const float2 *dc, hipStream_t st);
hipError_t launch_conv_rows_fp16(int log2n, bool corr, const void *x, void *y, long rows, const float2 *spec,
                                 const float2 *tw, float scale, hipStream_t st, const ConvGroup *group = nullptr);
hipError_t launch_conv_fp16_filter_group(ConvGroup *g, int members, hipStream_t st);
hipError_t launch_conv_scales_group(const ConvGroup *g, int members, long max_rows, long n, bool registered,
                                    hipStream_t st);
hipError_t launch_conv_fp16_filter(const float *h, long k, float2 *spec, long n, const float2 *dc, bool normalized,
                                   hipStream_t st);
hipError_t launch_conv_scales(const float *s_in, float *s_out, const float2 *l1, long rows, hipStream_t st);
hipError_t launch_conv_half_merge(float2 *zv, const float2 *spec, long n, bool corr, hipStream_t st,
                                  const ConvGroup *group = nullptr, int members = 0);
hipError_t launch_conv_ols(bool fp16, bool corr, int log2L, const void *x, void *y, long rows, int log2n, long k,
                           long oshift, const float2 *spec, const float2 *tw, const float *s_in, float *s_out,
                           hipStream_t st);

int device_cus() {
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return cache[dev] = cus;
}

// a single FP32 row long enough for global passes (n >= 2^14) runs at half length
bool conv_half_path(long n, long rows, int precision) {
    return precision == ESTL_PRECISION_FP32 && rows == 1 && n >= (1L << (kMaxRowsLog2 + 1));
}

// A filter whose k taps all lie within the first spectrum pass's stride n / B_1,
// unrotated, puts one nonzero value in each column of that pass, whose DFT is that value
// in every output: the pass is skipped and the second reads the taps (io_in = 4).
bool spectrum_skips_first_pass(const Plan &spec, long k, long shift) {
    if (spec.kind != Plan::PASSES || spec.steps.size() < 2 || shift != 0 || k < 1) return false;
    return k <= (spec.n >> spec.steps.front().log2b);
}

// The passes of a spectrum plan over the filter(s) given by `taps` (rin, fk, fshift, or a
// group descriptor): the first reads the taps (io_in = 3), or, when `skip`, the second
// does (io_in = 4); with `divide_l1` the last stores each group member's spectrum
// divided by its L (io_out = 4).  Destinations alternate back from `spec` with
// `scratch0`.
static int spectrum_passes(const Plan &plan, const PassArgs &taps, bool skip, bool divide_l1, float2 *spec,
                           float2 *scratch0, hipStream_t st) {
    const auto &steps = plan.steps;
    const int P = (int)steps.size();
    const int q0 = skip ? 1 : 0;
    const float2 *src = scratch0;  // not read by the first pass run (it reads the taps)
    for (int q = q0; q < P; ++q) {
        PassArgs a = steps[q].args;
        if (q == q0) {
            a.io_in = skip ? 4 : 3;
            a.rin = taps.rin;
            a.fk = taps.fk;
            a.fshift = taps.fshift;
        }
        a.group = taps.group;
        if (divide_l1 && q == P - 1) a.io_out = 4;
        float2 *dst = ((P - 1 - q) & 1) ? scratch0 : spec;
        hipError_t e = launch_pass(steps[q].log2b, steps[q].log2t, ESTL_FORWARD, src, dst, a, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "filter spectrum pass: %s", hipGetErrorString(e));
        src = dst;
    }
    return ESTL_OK;
}

int group_spectrum_plan(Plan *p, long n, long filters) {
    int rc = build_plan(p, n, filters, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
    if (rc) return rc;
    if (p->kind != Plan::PASSES) return set_error(ESTL_ERR_UNSUPPORTED, "group spectra need a multi-pass size");
    PassStep &s0 = p->steps.front();
    while ((1 << s0.log2t) > (1L << s0.args.log2S) && s0.log2t > 0) --s0.log2t;
    s0.args.io_in = 3;
    return ESTL_OK;
}

int conv_shape_init(ConvShape *c, long n, long rows, unsigned flags, int precision) {
    if (n < 2 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two >= 2", n);
    if (precision != ESTL_PRECISION_FP32 && precision != ESTL_PRECISION_FP16)
        return set_error(ESTL_ERR_UNSUPPORTED, "convolution precision %d not supported", precision);
    const bool h = precision == ESTL_PRECISION_FP16;
    c->precision = precision;
    if (rows < 1) return set_error(ESTL_ERR_INVALID, "rows=%ld must be >= 1", rows);
    const unsigned known = ESTL_CONV_CORRELATE | ESTL_CONV_NORMALIZE_KERNEL | ESTL_CONV_CENTER | ESTL_CONV_FULL_LENGTH;
    if (flags & ~known) return set_error(ESTL_ERR_INVALID, "unknown convolution flags 0x%x", flags);
    c->n = n;
    c->rows = rows;
    c->pairs = (rows + 1) / 2;
    c->flags = flags;
    c->corr = (flags & ESTL_CONV_CORRELATE) != 0;
    c->log2n = 0;
    while ((1L << c->log2n) < n) ++c->log2n;
    c->scale = 1.0f / (float)n;
    c->spec_bytes = sizeof(float2) * (size_t)(h ? n + 1 : n);
    int rc = build_plan(&c->spec_plan, n, 1, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
    if (rc) return rc;
    if (c->spec_plan.kind == Plan::PASSES) {
        // the first pass builds the filter from its taps through its direct load path
        PassStep &s0 = c->spec_plan.steps.front();
        while ((1 << s0.log2t) > (1L << s0.args.log2S) && s0.log2t > 0) --s0.log2t;
        s0.args.io_in = 3;
    } else if (c->spec_plan.kind != Plan::ROWS) {
        return set_error(ESTL_ERR_UNSUPPORTED, "convolution of n=%ld: no spectrum plan", n);
    } else {
        // k_filter_spectrum's table (kSpecTw): the rows plan's power table below 4096
        // points, the block table from 4096
        c->tw_spec = c->spec_plan.tw_rows;
        if (c->log2n >= kSpecBlockTwLog2 && (rc = block_table(current_device(), c->log2n, &c->tw_spec))) return rc;
    }
    if (c->log2n <= kMaxRowsLog2) {
        c->fused = true;
        c->scratch_bytes = 0;
        return power_table(current_device(), c->log2n, &c->tw_rows);  // k_conv_rows: twiddles by powers
    }
    for (int w = 0; w < 2; ++w) {
        if (c->log2n <= kOlsLog2 + w) break;
        if ((rc = block_table(current_device(), kOlsLog2 + w, &c->tw_ols[w]))) return rc;
        if ((rc = power_table(current_device(), kOlsLog2 + w, &c->tw_ols_pow[w]))) return rc;
    }
    // forward passes with the default split; the inverse runs the radices in the reverse
    // order, so the last forward pass and the first inverse pass cover the same columns
    // and fuse (with the spectrum multiply) into one middle pass (pass_tile.h, MID)
    // FP16 storage: every pass stores 1/sqrt(radix) of its result, so the forward and
    // inverse passes together scale by exactly 1/n
    Plan f, b;
    rc = build_plan(&f, n, c->pairs, ESTL_FORWARD, precision, 0, nullptr);
    if (rc) return rc;
    std::vector<int> rev;
    for (auto it = f.steps.rbegin(); it != f.steps.rend(); ++it) rev.push_back(it->log2b);
    rc = build_plan(&b, n, c->pairs, ESTL_BACKWARD, precision, 0, &rev);
    if (rc) return rc;
    const size_t P = f.steps.size();
    if (P < 2 || b.steps.size() != P) return set_error(ESTL_ERR_UNSUPPORTED, "convolution of n=%ld needs >= 2 passes", n);
    c->fused = false;
    c->fwd.assign(f.steps.begin(), f.steps.end() - 1);
    c->mid = f.steps.back();
    // the middle pass computes two radix-B DFTs and a spectrum multiply per tile: tiles
    // one column step narrower than a plain pass's (more workgroups per CU to overlap
    // the loads with the longer compute) measured 4-11 % faster per convolution over
    // 2^22 values (n = 2^15..2^20, FP32)
    if (!h && c->mid.log2t > 0) --c->mid.log2t;
    c->mid.args.corr = c->corr ? 1 : 0;
    if (h) c->mid.args.scale = f.steps.back().args.scale * b.steps.front().args.scale;
    c->inv.assign(b.steps.begin() + 1, b.steps.end());
    // the first pass reads real pairs through its direct (S >= T) load path
    PassStep &f0 = c->fwd.front();
    while ((1 << f0.log2t) > (1L << f0.args.log2S) && f0.log2t > 0) --f0.log2t;
    f0.args.io_in = 1;
    f0.args.real_rows = rows;
    c->inv.back().args.io_out = 2;
    c->inv.back().args.real_rows = rows;
    c->scratch_bytes = (h ? 4 : sizeof(float2)) * (size_t)n * (size_t)c->pairs;
    if (conv_half_path(n, rows, precision)) {
        // the half-length transforms (their scratch, n/2 values, fits the pair's)
        rc = build_plan(&c->half_fwd, n / 2, 1, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
        if (rc) return rc;
        rc = build_plan(&c->half_inv, n / 2, 1, ESTL_BACKWARD, ESTL_PRECISION_FP32, 0, nullptr);
        if (rc) return rc;
        c->half = true;
    }
    return ESTL_OK;
}

size_t conv_scratch_bytes(const ConvShape &c) {
    const size_t spec_need = c.spec_plan.kind == Plan::PASSES ? c.spec_plan.scratch_bytes : 0;
    return c.scratch_bytes > spec_need ? c.scratch_bytes : spec_need;
}

size_t conv_footprint(const ConvShape &c) {
    // spectrum + dc + two scratch buffers (the spectrum FFT of a multi-pass size borrows
    // the first)
    return c.spec_bytes + 256 + 2 * conv_scratch_bytes(c);
}

int conv_filter(const ConvShape &c, const ConvBuffers &b, const float *h, long k, hipStream_t st) {
    if (!h || k < 1) return set_error(ESTL_ERR_INVALID, "filter needs k >= 1 values");
    const long shift = (c.flags & ESTL_CONV_CENTER) ? ((k + 1) / 2) % c.n : 0;
    const bool normalize = (c.flags & ESTL_CONV_NORMALIZE_KERNEL) != 0;
    hipError_t e = hipSuccess;
    if (const int lw = b.ols) {
        // overlap-save: the unrotated filter's spectrum of the window length (the
        // rotation moves the outputs instead), normalised in the same launch; FP16
        // storage divides it by L and keeps (L, sum |h|) after the window's bins, in that
        // launch too
        const long L = 1L << lw;
        const bool h16 = c.precision == ESTL_PRECISION_FP16;
        e = launch_filter_spectrum(lw, h, k, 0, b.spec, c.tw_ols[lw - kOlsLog2], normalize, st, nullptr, 0, false,
                                   b.dc, h16 ? b.spec + L : nullptr);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "filter spectrum launch: %s", hipGetErrorString(e));
        return ESTL_OK;
    }
    if (c.spec_plan.kind == Plan::ROWS) {
        // fold, transform and normalise in one launch (the kernel sum to b.dc)
        e = launch_filter_spectrum(c.log2n, h, k, shift, b.spec, c.tw_spec, normalize, st, nullptr, 0,
                                   false, b.dc);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "filter spectrum launch: %s", hipGetErrorString(e));
    } else {
        // the spectrum plan's passes, the first run building the filter from its taps
        if (!b.scratch[0]) return set_error(ESTL_ERR_INVALID, "the filter spectrum of n=%ld needs scratch", c.n);
        PassArgs taps{};
        taps.rin = h;
        taps.fk = k;
        taps.fshift = shift;
        if (int rc = spectrum_passes(c.spec_plan, taps, spectrum_skips_first_pass(c.spec_plan, k, shift), false,
                                     b.spec, b.scratch[0], st))
            return rc;
    }
    if (normalize && c.spec_plan.kind != Plan::ROWS) {
        // divide by the kernel sum K^[0] (Convolution.h:141): spectrum / spectrum[0].re
        e = hipMemcpyAsync(b.dc, b.spec, sizeof(float2), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "spectrum dc copy: %s", hipGetErrorString(e));
        e = launch_spec_normalize(b.spec, c.n, b.dc, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "normalize launch: %s", hipGetErrorString(e));
    }
    if (c.precision == ESTL_PRECISION_FP16) {
        // divide the spectrum by L = sum |h| (/ |kernel sum| when normalised), kept at spec[n]
        e = launch_conv_fp16_filter(h, k, b.spec, c.n, b.dc, (c.flags & ESTL_CONV_NORMALIZE_KERNEL) != 0, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "FP16 filter scaling: %s", hipGetErrorString(e));
    }
    return ESTL_OK;
}

// The multi-pass chain of shape c: forward passes (the first reads the real pairs) ->
// middle pass (last forward pass, spectrum multiply, first inverse pass) -> inverse
// passes (the last writes the real pairs), ping-ponging between two scratch buffers.
// One request: rows x -> y with spectrum `spec`; a request group: every member's rows
// and spectrum from `group`, over `pairs` pairs (0: the shape's batch).  FP32 or
// binary16 storage by the shape's precision (FP32's last pass takes the 1/n scale here;
// binary16 passes carry their own 1/sqrt(radix) scaling).
static int conv_pass_chain(const ConvShape &c, const void *x, void *y, const float2 *spec, const ConvGroup *group,
                           long pairs, void *scratch0, void *scratch1, hipStream_t st) {
    const bool h = c.precision == ESTL_PRECISION_FP16;
    void *scr[2] = {scratch0, scratch1};
    int cur = 0;
    const void *src = scratch1;  // not read by the first pass (it reads the real pairs)
    auto run = [&](const PassStep &s, int dir) -> int {
        PassArgs a = s.args;
        if (h) {
            a.rin16 = x;
            a.rout16 = y;
        } else {
            a.rin = static_cast<const float *>(x);
            a.rout = static_cast<float *>(y);
            if (a.io_out == 2) a.scale = c.scale;
        }
        a.spectrum = spec;
        a.group = group;
        if (pairs > 0) a.batch = pairs;
        void *dst = scr[cur];
        hipError_t e;
        if (h)
            e = dir ? launch_pass_fp16(s.log2b, s.log2t, dir, src, dst, a, st)
                    : launch_pass_mid_fp16(s.log2b, s.log2t, src, dst, a, st);
        else
            e = dir ? launch_pass(s.log2b, s.log2t, dir, static_cast<const float2 *>(src), static_cast<float2 *>(dst), a,
                                  st)
                    : launch_pass_mid(s.log2b, s.log2t, static_cast<const float2 *>(src), static_cast<float2 *>(dst), a,
                                      st);
        if (e != hipSuccess)
            return set_error(ESTL_ERR_HIP, "%sconvolution pass launch: %s", h ? "FP16 " : "", hipGetErrorString(e));
        src = dst;
        cur ^= 1;
        return ESTL_OK;
    };
    for (const auto &s : c.fwd)
        if (int rc = run(s, ESTL_FORWARD)) return rc;
    if (int rc = run(c.mid, 0)) return rc;
    for (const auto &s : c.inv)
        if (int rc = run(s, ESTL_BACKWARD)) return rc;
    return ESTL_OK;
}

// One row through transforms of n/2 values (rows read and written as complex values,
// so both must be 8-byte aligned): forward plan into scratch0 (scratch1 its
// workspace), the merge in place, inverse plan into y.  The plans run with caller
// workspace, which leaves them unchanged.
static int conv_half_chain(const ConvShape &c, const float2 *spec, const float *x, float *y, float2 *scratch0,
                           float2 *scratch1, hipStream_t st) {
    if (!scratch0 || !scratch1) return set_error(ESTL_ERR_INVALID, "the convolution of n=%ld needs scratch", c.n);
    Plan *fwd = const_cast<Plan *>(&c.half_fwd), *inv = const_cast<Plan *>(&c.half_inv);
    if (int rc = execute_plan_ws(fwd, reinterpret_cast<const float2 *>(x), scratch0, scratch1, nullptr, st)) return rc;
    hipError_t e = launch_conv_half_merge(scratch0, spec, c.n, c.corr, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "k_conv_half_merge launch: %s", hipGetErrorString(e));
    return execute_plan_ws(inv, scratch0, reinterpret_cast<float2 *>(y), scratch1, nullptr, st);
}

// the overlap-save launch of a request whose filter (b.k taps, spectrum of 2^b.ols bins)
// takes that path
static int conv_run_ols(const ConvShape &c, const ConvBuffers &b, const void *x, void *y, const float *s_in,
                        float *s_out, hipStream_t st) {
    const bool fp16 = s_out != nullptr;
    const long shift = (c.flags & ESTL_CONV_CENTER) ? ((b.k + 1) / 2) % c.n : 0;
    // the filter rotation of ESTL_CONV_CENTER as an output rotation: a convolution's
    // result moves back by the shift, a correlation's forward
    const long oshift = c.corr ? shift : -shift;
    const int lw = b.ols;
    if (lw != kOlsLog2 && !(lw == kOlsWideLog2 && c.tw_ols_pow[1]))
        return set_error(ESTL_ERR_INVALID, "no overlap-save window of 2^%d values for n=%ld", lw, c.n);
    hipError_t e = launch_conv_ols(fp16, c.corr, lw, x, y, c.rows, c.log2n, b.k, oshift, b.spec,
                                   c.tw_ols_pow[lw - kOlsLog2], s_in, s_out, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "k_conv_ols launch: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int conv_run(const ConvShape &c, const ConvBuffers &b, const float *x, float *y, hipStream_t st) {
    if (!x || !y) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (c.precision != ESTL_PRECISION_FP32)
        return set_error(ESTL_ERR_INVALID, "an FP16-storage convolution runs through estl_conv_execute_fp16");
    if (b.ols) return conv_run_ols(c, b, x, y, nullptr, nullptr, st);
    if (c.half && !(reinterpret_cast<uintptr_t>(x) & 7) && !(reinterpret_cast<uintptr_t>(y) & 7))
        return conv_half_chain(c, b.spec, x, y, b.scratch[0], b.scratch[1], st);
    if (c.fused) {
        hipError_t e = launch_conv_rows(c.log2n, c.corr, x, y, c.rows, b.spec, c.tw_rows, c.scale, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "k_conv_rows launch: %s", hipGetErrorString(e));
        return ESTL_OK;
    }
    return conv_pass_chain(c, x, y, b.spec, nullptr, 0, b.scratch[0], b.scratch[1], st);
}

// A request group (ConvGroup in device memory at d_group).  Spectra: the G filters as
// the G rows of `spec` (a spectrum plan over G rows whose first pass builds them from
// their taps), destinations alternating back from `spectra` with scratch0.
int conv_filter_group_rows(const ConvShape &c, const ConvGroup *d_group, int members, bool divide_l1,
                           hipStream_t st) {
    if (!c.fused || c.spec_plan.kind != Plan::ROWS) return set_error(ESTL_ERR_INVALID, "not a fused shape");
    hipError_t e = launch_filter_spectrum(c.log2n, nullptr, 0, 0, nullptr, c.tw_spec, false, st, d_group,
                                          members, divide_l1);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group filter spectra: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int conv_filter_group(const Plan &spec, const ConvGroup *d_group, float2 *spectra, float2 *scratch0,
                      bool skip_first, bool divide_l1, hipStream_t st) {
    if (spec.kind != Plan::PASSES || spec.steps.empty() || spec.steps.front().args.io_in != 3)
        return set_error(ESTL_ERR_INVALID, "not a group spectrum plan");
    if (skip_first && spec.steps.size() < 2) return set_error(ESTL_ERR_INVALID, "no first pass to skip");
    PassArgs taps{};
    taps.group = d_group;
    return spectrum_passes(spec, taps, skip_first, divide_l1, spectra, scratch0, st);
}

// The convolution passes of a request group over its `pairs` pairs (the members' pairs
// back to back): the first pass reads each member's rows, the middle pass multiplies by
// each member's spectrum, the last writes each member's rows.  The shape may have been
// built for another batch: a pass's arithmetic and tiles do not depend on it, only the
// grid, which follows `pairs`.
int conv_run_group(const ConvShape &c, const ConvGroup *d_group, long pairs, float2 *scratch0, float2 *scratch1,
                   hipStream_t st) {
    if (c.precision != ESTL_PRECISION_FP32) return set_error(ESTL_ERR_INVALID, "request groups run FP32 convolutions");
    if (pairs < 1) return set_error(ESTL_ERR_INVALID, "empty request group");
    if (c.fused) {  // one launch over the group's pairs
        hipError_t e = launch_conv_rows(c.log2n, c.corr, nullptr, nullptr, 2 * pairs, nullptr, c.tw_rows, c.scale, st,
                                        d_group);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group k_conv_rows launch: %s", hipGetErrorString(e));
        return ESTL_OK;
    }
    return conv_pass_chain(c, nullptr, nullptr, nullptr, d_group, pairs, scratch0, scratch1, st);
}

// A request group of single FP32 rows (conv_half_path, n/2 > 2^13): the forward
// passes of n/2 over the members' rows (the first reading each member's row as
// complex values) into z, the merge per member, the inverse passes (the last writing
// each member's output), ping-ponging with ws.  fwd and inv are plans of n/2 values
// for any batch: every pass runs over `members` rows.
int conv_run_group_half(const Plan &fwd, const Plan &inv, const ConvGroup *d_group, int members, long n, bool corr,
                        float2 *z, float2 *ws, hipStream_t st) {
    if (fwd.kind != Plan::PASSES || inv.kind != Plan::PASSES || fwd.n * 2 != n || inv.n * 2 != n ||
        fwd.steps.size() != inv.steps.size() || fwd.steps.size() < 2)
        return set_error(ESTL_ERR_INVALID, "not a half-length group plan pair for n=%ld", n);
    if (members < 1 || members > kGroupMax || !z || !ws) return set_error(ESTL_ERR_INVALID, "bad request group");
    const int P = (int)fwd.steps.size();
    const float2 *src = nullptr;  // the first pass reads the members' rows
    for (int q = 0; q < P; ++q) {
        PassArgs a = fwd.steps[q].args;
        a.batch = members;
        if (q == 0) {
            a.io_in = 5;
            a.group = d_group;
        }
        float2 *dst = ((P - 1 - q) & 1) ? ws : z;
        hipError_t e = launch_pass(fwd.steps[q].log2b, fwd.steps[q].log2t, ESTL_FORWARD, src, dst, a, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group half-length forward pass: %s", hipGetErrorString(e));
        src = dst;
    }
    hipError_t e = launch_conv_half_merge(z, nullptr, n, corr, st, d_group, members);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group k_conv_half_merge launch: %s", hipGetErrorString(e));
    src = z;
    for (int q = 0; q < P; ++q) {
        PassArgs a = inv.steps[q].args;
        a.batch = members;
        float2 *dst = (q & 1) ? z : ws;
        if (q == P - 1) {
            a.io_out = 3;
            a.group = d_group;
            dst = nullptr;  // the last pass writes the members' outputs
        }
        e = launch_pass(inv.steps[q].log2b, inv.steps[q].log2t, ESTL_BACKWARD, src, dst, a, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group half-length inverse pass: %s", hipGetErrorString(e));
        src = dst;
    }
    return ESTL_OK;
}

// FP16 storage: every member's spectrum divided by its filter's L = sum |h|.  Inline
// filters: L into the descriptor before the spectra (conv_fp16_filter_group), whose
// last launch stores them divided by it (divide_l1); a registered filter's spectrum
// was divided when it was registered, with L after its n values.  Then the passes on
// binary16 rows and the members' output row scales (conv_run_group_fp16).
int conv_fp16_filter_group(ConvGroup *d_group, int members, hipStream_t st) {
    hipError_t e = launch_conv_fp16_filter_group(d_group, members, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group FP16 filter scaling: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int conv_run_group_fp16(const ConvShape &c, ConvGroup *d_group, int members, long pairs, long max_rows,
                        bool registered, void *scratch0, void *scratch1, hipStream_t st) {
    if (c.precision != ESTL_PRECISION_FP16) return set_error(ESTL_ERR_INVALID, "not an FP16-storage shape");
    if (pairs < 1) return set_error(ESTL_ERR_INVALID, "empty request group");
    hipError_t e;
    if (c.fused) {
        e = launch_conv_rows_fp16(c.log2n, c.corr, nullptr, nullptr, 2 * pairs, nullptr, c.tw_rows, c.scale, st,
                                  d_group);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group FP16 k_conv_rows launch: %s", hipGetErrorString(e));
    } else if (int rc = conv_pass_chain(c, nullptr, nullptr, nullptr, d_group, pairs, scratch0, scratch1, st)) {
        return rc;
    }
    e = launch_conv_scales_group(d_group, members, max_rows, c.n, registered, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "group output scales: %s", hipGetErrorString(e));
    return ESTL_OK;
}

int conv_run_fp16(const ConvShape &c, const ConvBuffers &b, const void *x16, const float *s_in, void *y16,
                  float *s_out, hipStream_t st) {
    if (!x16 || !y16 || !s_in || !s_out) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (c.precision != ESTL_PRECISION_FP16) return set_error(ESTL_ERR_INVALID, "not an FP16-storage convolution");
    hipError_t e;
    if (b.ols) {
        return conv_run_ols(c, b, x16, y16, s_in, s_out, st);  // the output scales in the same launch
    }
    if (c.fused) {
        e = launch_conv_rows_fp16(c.log2n, c.corr, x16, y16, c.rows, b.spec, c.tw_rows, c.scale, st);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "FP16 k_conv_rows launch: %s", hipGetErrorString(e));
    } else if (int rc = conv_pass_chain(c, x16, y16, b.spec, nullptr, 0, b.scratch[0], b.scratch[1], st)) {
        return rc;
    }
    e = launch_conv_scales(s_in, s_out, b.spec + c.n, c.rows, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "output scales: %s", hipGetErrorString(e));
    return ESTL_OK;
}

// the public object: a shape plus buffers of its own
struct Conv {
    ConvShape shape;
    ConvBuffers buf;
    bool have_filter = false;
};

static void conv_free(Conv *c) {
    if (!c) return;
    for (float2 *p : {c->buf.spec, c->buf.dc, c->buf.scratch[0], c->buf.scratch[1]})
        if (p) (void)hipFree(p);
    free_stream_scratch(&c->shape.spec_plan);
    for (auto &s : c->shape.spec_plan.scratch)
        if (s) (void)hipFree(s);
    delete c;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_conv_create(estl_conv_t *conv, long n, long rows, int precision, unsigned flags) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle pointer is null");
    *conv = nullptr;
    Conv *c = new Conv();
    int rc = conv_shape_init(&c->shape, n, rows, flags, precision);
    if (rc) {
        conv_free(c);
        return rc;
    }
    const ConvShape &s = c->shape;
    bool ok = hipMalloc(&c->buf.spec, s.spec_bytes) == hipSuccess && hipMalloc(&c->buf.dc, sizeof(float2)) == hipSuccess;
    const size_t sb = conv_scratch_bytes(s);
    for (auto &p : c->buf.scratch)
        if (ok && sb) ok = hipMalloc(&p, sb) == hipSuccess;
    if (!ok) {
        conv_free(c);
        return set_error(ESTL_ERR_NOMEM, "convolution buffers (n=%ld, rows=%ld)", n, rows);
    }
    *conv = reinterpret_cast<estl_conv_t>(c);
    return ESTL_OK;
}

int estl_conv_set_filter(estl_conv_t conv, const float *d_h, long k, void *stream) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    Conv *c = reinterpret_cast<Conv *>(conv);
    if (!d_h || k < 1) return set_error(ESTL_ERR_INVALID, "filter needs k >= 1 values");  // the old filter stays
    const ConvShape &s = c->shape;
    // a launch of its own: the window length by the device's CU count
    ConvBuffers b = c->buf;
    b.k = k;
    b.ols = conv_ols_log2(s.n, s.rows, k, s.flags, device_cus());
    int rc = conv_filter(s, b, d_h, k, static_cast<hipStream_t>(stream));
    c->have_filter = rc == ESTL_OK;  // a failed call may have overwritten the spectrum
    c->buf = b;
    return rc;
}

int estl_conv_execute(estl_conv_t conv, const float *d_x, float *d_y, void *stream) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    Conv *c = reinterpret_cast<Conv *>(conv);
    if (!c->have_filter) return set_error(ESTL_ERR_INVALID, "estl_conv_set_filter has not been called");
    return conv_run(c->shape, c->buf, d_x, d_y, static_cast<hipStream_t>(stream));
}

int estl_conv_execute_fp16(estl_conv_t conv, const void *d_x16, const float *d_x_scale, void *d_y16,
                           float *d_y_scale, void *stream) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    Conv *c = reinterpret_cast<Conv *>(conv);
    if (!c->have_filter) return set_error(ESTL_ERR_INVALID, "estl_conv_set_filter has not been called");
    return conv_run_fp16(c->shape, c->buf, d_x16, d_x_scale, d_y16, d_y_scale, static_cast<hipStream_t>(stream));
}

long long estl_conv_workspace_bytes(estl_conv_t conv) {
    if (!conv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return (long long)conv_footprint(reinterpret_cast<Conv *>(conv)->shape);
}

int estl_conv_destroy(estl_conv_t conv) {
    conv_free(reinterpret_cast<Conv *>(conv));
    return ESTL_OK;
}

}  // extern "C"
