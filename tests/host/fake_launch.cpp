// fake_launch.cpp -- CPU stand-ins for the library's plan and launch functions, so the
// request server (estl_amd/csrc/server.cpp, compiled unchanged against the fake HIP
// runtime in fake_hip/) runs on the CPU under ThreadSanitizer / AddressSanitizer.
//
// Test infrastructure only: nothing here is built into libestl_hip.so.  Each "launch"
// is a task on the fake stream; the arithmetic is plain double-precision FFTs, so a
// request's result is a deterministic function of its inputs, whichever code path (alone
// or in a request group) produced it.  What is modelled faithfully is the data flow the
// server must order: which buffers each launch reads and writes (outputs, the two
// scratch buffers, spectra, the group descriptor read at run time from device memory)
// and when (in stream order, after the launches queued before it).
//
// Fault injection (fake_set_fail_launch / fake_set_fail_async): the k-th launch counted
// from the last reset fails when it is issued (ESTL_ERR_HIP, nothing queued), or is
// queued and fails when it runs (its stream's next event reports hipErrorLaunchFailure).
// fake_set_stall_us delays every task, so launches queued before a failure are still
// pending when the server reacts to it.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/estl_hip.h"
#include "../../estl_amd/csrc/plan.h"

namespace {

thread_local std::string g_err;
std::atomic<long> g_launches{0};
std::atomic<long> g_fail_sync{0}, g_fail_async{0};

using cd = std::complex<double>;

void fft(std::vector<cd> &a, int dir) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const double ang = dir * 2.0 * M_PI * (double)k / (double)len;
                const cd w(std::cos(ang), std::sin(ang));
                const cd u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// the filter's spectrum: taps folded modulo n and rotated by `shift`, transformed
void filter_spectrum(const float *h, long k, long shift, long n, bool normalize, float2 *spec) {
    std::vector<cd> f(n);
    for (long i = 0; i < k; ++i) f[(size_t)(((i - shift) % n + n) % n)] += h[i];
    fft(f, -1);
    const double dc = normalize ? f[0].real() : 1.0;
    for (long i = 0; i < n; ++i) spec[i] = {(float)(f[i].real() / dc), (float)(f[i].imag() / dc)};
}

float l1_of(const float *h, long k) {
    double s = 0;
    for (long i = 0; i < k; ++i) s += std::fabs((double)h[i]);
    return (float)s;
}

void conv_row(const float *x, long n, const float2 *spec, bool corr, float *y) {
    std::vector<cd> a(n);
    for (long i = 0; i < n; ++i) a[i] = x[i];
    fft(a, -1);
    for (long i = 0; i < n; ++i) a[i] *= corr ? cd(spec[i].x, -spec[i].y) : cd(spec[i].x, spec[i].y);
    fft(a, +1);
    for (long i = 0; i < n; ++i) y[i] = (float)(a[i].real() / (double)n);
}

// FP16 storage stand-in: the binary16 rows pass through unchanged, the output row scale
// is the input's times the filter's L
void conv_rows16(const void *x16, const float *s_in, long n, long rows, float L, void *y16, float *s_out) {
    std::memcpy(y16, x16, sizeof(uint16_t) * (size_t)n * (size_t)rows);
    for (long r = 0; r < rows; ++r) s_out[r] = s_in[r] * L;
}

// Count a launch: fail it now, queue it as failing, or queue `work` on st.
int launch(hipStream_t st, const char *what, std::function<void()> work) {
    const long id = ++g_launches;
    if (id == g_fail_sync.load()) return estl::set_error(ESTL_ERR_HIP, "injected failure of launch %ld (%s)", id, what);
    if (id == g_fail_async.load()) {
        fakehip::enqueue(st, [st] {
            if (st) {
                std::lock_guard<std::mutex> lk(st->mu);
                st->failed = true;
            }
        });
        return ESTL_OK;
    }
    fakehip::enqueue(st, std::move(work));
    return ESTL_OK;
}

int ilog2(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

}  // namespace

extern "C" {
void fake_set_fail_launch(long k) { g_fail_sync = k; }
void fake_set_fail_async(long k) { g_fail_async = k; }
void fake_reset_launches() { g_launches = 0; }
long fake_launches() { return g_launches.load(); }
void fake_set_stall_us(int us) { fakehip::stall_us() = us; }
void fake_stall_next_task_us(int us) { fakehip::stall_once_us() = us; }
const char *estl_last_error(void) { return g_err.c_str(); }
}

namespace estl {

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int current_device() { return 0; }

int build_plan(Plan *p, long n, long batch, int dir, int precision, unsigned flags, const std::vector<int> *split,
               const std::vector<int> *) {
    if (n < 1 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two", n);
    if (batch < 1) return set_error(ESTL_ERR_INVALID, "batch=%ld must be >= 1", batch);
    if (dir != ESTL_FORWARD && dir != ESTL_BACKWARD) return set_error(ESTL_ERR_INVALID, "bad direction %d", dir);
    if (precision != ESTL_PRECISION_FP32 && precision != ESTL_PRECISION_FP16)
        return set_error(ESTL_ERR_UNSUPPORTED, "precision %d", precision);
    p->n = n;
    p->batch = batch;
    p->dir = dir;
    p->precision = precision;
    p->flags = flags;
    p->log2n = ilog2(n);
    p->scale = (flags & ESTL_FLAG_NORMALIZE) ? 1.0f / (float)n : 1.0f;
    p->steps.clear();
    if (n == 1) {
        p->kind = Plan::COPY;
    } else if (p->log2n <= kMaxRowsLog2 && !(split && !split->empty())) {
        p->kind = Plan::ROWS;
    } else {
        p->kind = Plan::PASSES;
        std::vector<int> bits = split && !split->empty() ? *split : std::vector<int>();
        if (bits.empty()) {
            const int P = std::max(2, (p->log2n + 8) / 9);
            for (int q = 0; q < P; ++q) bits.push_back(p->log2n / P + (q < p->log2n % P ? 1 : 0));
        }
        for (int b : bits) {
            PassStep s{};
            s.log2b = b;
            p->steps.push_back(s);
        }
        p->scratch_bytes = (precision == ESTL_PRECISION_FP16 ? 4 : sizeof(float2)) * (size_t)n * (size_t)batch;
    }
    return ESTL_OK;
}

// the plan's transform of every row of `in` into `out` (FP32 plans)
static void transform_rows(const Plan &p, const float2 *in, float2 *out) {
    std::vector<cd> a(p.n);
    for (long r = 0; r < p.batch; ++r) {
        for (long i = 0; i < p.n; ++i) a[i] = cd(in[r * p.n + i].x, in[r * p.n + i].y);
        fft(a, p.dir);
        for (long i = 0; i < p.n; ++i)
            out[r * p.n + i] = {(float)(a[i].real() * p.scale), (float)(a[i].imag() * p.scale)};
    }
}

int execute_plan_ws(Plan *p, const float2 *in, float2 *out, float2 *ws0, float2 *ws1, hipStream_t st) {
    if (!in || !out) return set_error(ESTL_ERR_INVALID, "null buffer");
    const Plan *pp = p;
    const size_t bytes = sizeof(float2) * (size_t)p->n * (size_t)p->batch;
    switch (p->kind) {
        case Plan::COPY:
            return launch(st, "copy", [=] { std::memmove(out, in, bytes); });
        case Plan::ROWS:
            return launch(st, "rows", [=] { transform_rows(*pp, in, out); });
        case Plan::PASSES: {
            if (!ws0) return set_error(ESTL_ERR_INVALID, "the fake passes need caller workspace");
            const int P = (int)p->steps.size();
            std::vector<float2 *> dst(P);
            for (int q = P - 1, k = 0; q >= 0; --q, ++k) dst[q] = (k & 1) ? ws0 : out;
            if (P > 1 && dst[0] == in) {
                if (!ws1) return set_error(ESTL_ERR_INVALID, "an in-place odd-length chain needs a second workspace");
                dst[0] = ws1;
            }
            // the first pass computes the transform, the others move it along the chain
            if (int rc = launch(st, "pass 0", [=] { transform_rows(*pp, in, dst[0]); })) return rc;
            for (int q = 1; q < P; ++q) {
                float2 *s = dst[q - 1], *d = dst[q];
                if (int rc = launch(st, "pass", [=] { std::memmove(d, s, bytes); })) return rc;
            }
            return ESTL_OK;
        }
    }
    return set_error(ESTL_ERR_INVALID, "bad plan");
}

bool conv_half_path(long n, long rows, int precision) {
    return precision == ESTL_PRECISION_FP32 && rows == 1 && n >= (1L << (kMaxRowsLog2 + 1));
}

bool spectrum_skips_first_pass(const Plan &spec, long k, long shift) {
    if (spec.kind != Plan::PASSES || spec.steps.size() < 2 || shift != 0 || k < 1) return false;
    return k <= (spec.n >> spec.steps.front().log2b);
}

int group_spectrum_plan(Plan *p, long n, long filters) {
    int rc = build_plan(p, n, filters, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
    if (rc) return rc;
    if (p->kind != Plan::PASSES) return set_error(ESTL_ERR_UNSUPPORTED, "group spectra need a multi-pass size");
    return ESTL_OK;
}

int conv_shape_init(ConvShape *c, long n, long rows, unsigned flags, int precision) {
    if (n < 2 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two >= 2", n);
    if (precision != ESTL_PRECISION_FP32 && precision != ESTL_PRECISION_FP16)
        return set_error(ESTL_ERR_UNSUPPORTED, "convolution precision %d not supported", precision);
    if (rows < 1) return set_error(ESTL_ERR_INVALID, "rows=%ld must be >= 1", rows);
    const unsigned known = ESTL_CONV_CORRELATE | ESTL_CONV_NORMALIZE_KERNEL | ESTL_CONV_CENTER | ESTL_CONV_FULL_LENGTH;
    if (flags & ~known) return set_error(ESTL_ERR_INVALID, "unknown convolution flags 0x%x", flags);
    const bool h = precision == ESTL_PRECISION_FP16;
    c->precision = precision;
    c->n = n;
    c->rows = rows;
    c->pairs = (rows + 1) / 2;
    c->flags = flags;
    c->corr = (flags & ESTL_CONV_CORRELATE) != 0;
    c->log2n = ilog2(n);
    c->scale = 1.0f / (float)n;
    c->spec_bytes = sizeof(float2) * (size_t)(h ? n + 1 : n);
    int rc = build_plan(&c->spec_plan, n, 1, ESTL_FORWARD, ESTL_PRECISION_FP32, 0, nullptr);
    if (rc) return rc;
    c->fused = c->log2n <= kMaxRowsLog2;
    c->scratch_bytes = c->fused ? 0 : (h ? 4 : sizeof(float2)) * (size_t)n * (size_t)c->pairs;
    c->half = !c->fused && conv_half_path(n, rows, precision);
    return ESTL_OK;
}

int device_cus() { return 256; }

size_t conv_scratch_bytes(const ConvShape &c) {
    const size_t spec_need = c.spec_plan.kind == Plan::PASSES ? c.spec_plan.scratch_bytes : 0;
    return c.scratch_bytes > spec_need ? c.scratch_bytes : spec_need;
}

// The overlap-save path (conv_ols_log2) keeps a spectrum of the window length, 2^13 or
// 2^14 values (plus the FP16 L1 pair): here the window holds the filter's taps themselves,
// (h_i, 0) zero-padded, and the convolution builds the n-point spectrum from them when it
// runs -- the same buffer sizes and the same data flow as the library's.
static void ols_taps(const float *h, long k, long L, float2 *spec) {
    for (long i = 0; i < L; ++i) spec[i] = {i < k ? h[i] : 0.0f, 0.0f};
}
static void ols_spectrum(const ConvShape &c, const float2 *taps, long k, std::vector<float2> &spec) {
    std::vector<float> h(k);
    for (long i = 0; i < k; ++i) h[i] = taps[i].x;
    const long shift = (c.flags & ESTL_CONV_CENTER) ? ((k + 1) / 2) % c.n : 0;
    spec.resize(c.n);
    filter_spectrum(h.data(), k, shift, c.n, (c.flags & ESTL_CONV_NORMALIZE_KERNEL) != 0, spec.data());
}

int conv_filter(const ConvShape &c, const ConvBuffers &b, const float *h, long k, hipStream_t st) {
    if (!h || k < 1) return set_error(ESTL_ERR_INVALID, "filter needs k >= 1 values");
    const long shift = (c.flags & ESTL_CONV_CENTER) ? ((k + 1) / 2) % c.n : 0;
    const bool normalize = (c.flags & ESTL_CONV_NORMALIZE_KERNEL) != 0;
    const bool fp16 = c.precision == ESTL_PRECISION_FP16;
    const long n = c.n;
    float2 *spec = b.spec, *scr = b.scratch[0];
    if (b.ols) {
        const long L = 1L << b.ols;
        return launch(st, "ols filter", [=] {
            ols_taps(h, k, L, spec);
            if (fp16) spec[L] = {l1_of(h, k), l1_of(h, k)};
        });
    }
    if (c.spec_plan.kind == Plan::PASSES) {
        if (!scr) return set_error(ESTL_ERR_INVALID, "the filter spectrum of n=%ld needs scratch", n);
        // first pass into scratch 0, the last into the spectrum
        if (int rc = launch(st, "spectrum pass 0", [=] { filter_spectrum(h, k, shift, n, normalize, scr); })) return rc;
        if (int rc = launch(st, "spectrum pass", [=] { std::memcpy(spec, scr, sizeof(float2) * (size_t)n); }))
            return rc;
    } else {
        if (int rc = launch(st, "filter spectrum", [=] { filter_spectrum(h, k, shift, n, normalize, spec); }))
            return rc;
    }
    if (fp16) {
        if (int rc = launch(st, "fp16 filter", [=] {
                const float L = l1_of(h, k);
                for (long i = 0; i < n; ++i) spec[i] = {spec[i].x / L, spec[i].y / L};
                spec[n] = {L, L};
            }))
            return rc;
    }
    return ESTL_OK;
}

int conv_run(const ConvShape &c, const ConvBuffers &b, const float *x, float *y, hipStream_t st) {
    const long n = c.n, rows = c.rows;
    const bool corr = c.corr;
    const float2 *spec = b.spec;
    if (b.ols) {
        const ConvShape *cp = &c;
        const long k = b.k;
        return launch(st, "conv ols", [=] {
            std::vector<float2> s;
            ols_spectrum(*cp, spec, k, s);
            for (long r = 0; r < rows; ++r) conv_row(x + r * n, n, s.data(), corr, y + r * n);
        });
    }
    if (c.fused)
        return launch(st, "conv rows", [=] {
            for (long r = 0; r < rows; ++r) conv_row(x + r * n, n, spec, corr, y + r * n);
        });
    float *s0 = reinterpret_cast<float *>(b.scratch[0]), *s1 = reinterpret_cast<float *>(b.scratch[1]);
    if (!s0 || !s1) return set_error(ESTL_ERR_INVALID, "convolution passes need two scratch buffers");
    const size_t bytes = sizeof(float) * (size_t)n * (size_t)rows;
    if (int rc = launch(st, "conv forward", [=] {
            for (long r = 0; r < rows; ++r) conv_row(x + r * n, n, spec, corr, s0 + r * n);
        }))
        return rc;
    if (int rc = launch(st, "conv middle", [=] { std::memcpy(s1, s0, bytes); })) return rc;
    return launch(st, "conv inverse", [=] { std::memcpy(y, s1, bytes); });
}

int conv_run_fp16(const ConvShape &c, const ConvBuffers &b, const void *x16, const float *s_in, void *y16, float *s_out,
                  hipStream_t st) {
    const long n = c.n, rows = c.rows;
    const float2 *spec = b.spec;
    const long l1_at = b.ols ? (1L << b.ols) : n;
    return launch(st, "conv fp16", [=] { conv_rows16(x16, s_in, n, rows, spec[l1_at].y, y16, s_out); });
}

// a group descriptor as the kernels see it: read from device memory when the launch runs
static ConvGroup read_group(const ConvGroup *d) {
    ConvGroup g;
    std::memcpy(&g, d, sizeof g);
    return g;
}

int conv_filter_group(const Plan &spec, const ConvGroup *d_group, float2 *spectra, float2 *scratch0, bool,
                      bool divide_l1, hipStream_t st) {
    const long n = spec.n;
    (void)spectra;
    if (int rc = launch(st, "group spectra pass 0", [=] {
            const ConvGroup g = read_group(d_group);
            for (int q = 0; q < g.count; ++q) filter_spectrum(g.h[q], g.k[q], g.shift[q], n, false, scratch0 + q * n);
        }))
        return rc;
    return launch(st, "group spectra pass", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q) {
            float2 *dst = const_cast<float2 *>(g.spec[q]);
            const float L = divide_l1 ? g.l1[q].y : 1.0f;
            for (long i = 0; i < n; ++i) dst[i] = {scratch0[q * n + i].x / L, scratch0[q * n + i].y / L};
        }
    });
}

int conv_filter_group_rows(const ConvShape &c, const ConvGroup *d_group, int, bool divide_l1, hipStream_t st) {
    const long n = c.n;
    return launch(st, "group filter rows", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q) {
            float2 *dst = const_cast<float2 *>(g.spec[q]);
            filter_spectrum(g.h[q], g.k[q], g.shift[q], n, false, dst);
            if (divide_l1) {
                const float L = g.l1[q].y;
                for (long i = 0; i < n; ++i) dst[i] = {dst[i].x / L, dst[i].y / L};
            }
        }
    });
}

int conv_fp16_filter_group(ConvGroup *d_group, int, hipStream_t st) {
    return launch(st, "group fp16 L", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q) {
            const float L = l1_of(g.h[q], g.k[q]);
            d_group->l1[q] = {L, L};
        }
    });
}

int conv_run_group(const ConvShape &c, const ConvGroup *d_group, long pairs, float2 *scratch0, float2 *scratch1,
                   hipStream_t st) {
    const long n = c.n;
    const bool corr = c.corr;
    auto body = [=](bool through_scratch) {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q) {
            const float *x = static_cast<const float *>(g.x[q]);
            float *y = through_scratch ? reinterpret_cast<float *>(scratch0) + 2 * n * g.pair_start[q]
                                       : static_cast<float *>(g.y[q]);
            for (long r = 0; r < g.real_rows[q]; ++r) conv_row(x + r * n, n, g.spec[q], corr, y + r * n);
        }
    };
    if (c.fused) return launch(st, "group conv rows", [=] { body(false); });
    if (int rc = launch(st, "group conv forward", [=] { body(true); })) return rc;
    const size_t bytes = sizeof(float2) * (size_t)n * (size_t)pairs;
    if (int rc = launch(st, "group conv middle", [=] { std::memcpy(scratch1, scratch0, bytes); })) return rc;
    return launch(st, "group conv inverse", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q)
            std::memcpy(g.y[q], reinterpret_cast<const float *>(scratch1) + 2 * n * g.pair_start[q],
                        sizeof(float) * (size_t)n * (size_t)g.real_rows[q]);
    });
}

int conv_run_group_fp16(const ConvShape &c, ConvGroup *d_group, int, long, long, bool registered, void *, void *,
                        hipStream_t st) {
    const long n = c.n;
    return launch(st, "group conv fp16", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q) {
            const float L = registered ? g.spec[q][n].y : g.l1[q].y;
            conv_rows16(g.x[q], g.xscale[q], n, g.real_rows[q], L, g.y[q], g.yscale[q]);
        }
    });
}

int conv_run_group_half(const Plan &, const Plan &, const ConvGroup *d_group, int, long n, bool corr, float2 *z,
                        float2 *ws, hipStream_t st) {
    if (int rc = launch(st, "group half forward", [=] {
            const ConvGroup g = read_group(d_group);
            for (int q = 0; q < g.count; ++q)
                conv_row(static_cast<const float *>(g.x[q]), n, g.spec[q], corr, reinterpret_cast<float *>(z) + q * n);
        }))
        return rc;
    if (int rc = launch(st, "group half merge", [=] {
            const ConvGroup g = read_group(d_group);
            std::memcpy(ws, z, sizeof(float) * (size_t)n * (size_t)g.count);
        }))
        return rc;
    return launch(st, "group half inverse", [=] {
        const ConvGroup g = read_group(d_group);
        for (int q = 0; q < g.count; ++q)
            std::memcpy(g.y[q], reinterpret_cast<const float *>(ws) + q * n, sizeof(float) * (size_t)n);
    });
}

}  // namespace estl
