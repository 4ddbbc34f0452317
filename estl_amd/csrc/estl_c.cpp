// estl_c.cpp -- host side of the C ABI declared in include/estl_hip.h: plans, the
// per-device twiddle cache, execution and status reporting.
//
// The planner is the counterpart of FFT_OpenCL_Contiguous's constructor
// (FFT_OpenCL_Contiguous.h:46-71): where the reference emits one OpenCL kernel string
// per radix-16 pass and compiles a program per FFT object, a plan here picks among
// precompiled gfx950 kernels:
//   * rows plan   (n <= 2^13): one launch of k_fft_rows, all passes in LDS;
//   * pass plan   (n > 2^13): P launches of k_fft_pass, radix up to 2^13 each, the
//                 Stockham pass formula of FFT_OpenCL_Contiguous.h:167-180 with its
//                 A_q / N_q prefix products (computeParameters, OpenCLFFTAlgorithm.h:9-12).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <cmath>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}
const char *last_error() { return g_last_error.c_str(); }



// ------------------------------------------------------------------ device state
static thread_local int g_device = -1;

int current_device() {
    if (g_device < 0) {
        int d = 0;
        if (hipGetDevice(&d) != hipSuccess) d = 0;
        g_device = d;
    }
    return g_device;
}

// Twiddle tables, built once per device and key, resident until process exit.
struct TwiddleCache {
    std::mutex mu;
    std::map<std::tuple<int, int, long, long, long>, float2 *> tables;  // (device, kind, a, b, c)
};
static TwiddleCache &twcache() {
    static TwiddleCache *c = new TwiddleCache;  // never destroyed: tables live to process exit
    return *c;
}

// Tables are generated on a private stream and waited for on that stream only, so a
// size seen for the first time does not stall work already running on the device.
static hipStream_t table_stream(int device) {
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    std::lock_guard<std::mutex> lk(mu);
    auto it = streams.find(device);
    if (it != streams.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    streams[device] = s;
    return s;
}

// Power table of BlockGeom<log2n> (block_fft's kTwPowers): W_{N_q}^g at powoff(q) + g,
// g in [0, A_q), per pass q >= 1; FP64 on the device, rounded once.
int power_table(int device, int log2n, const float2 **out) {
    auto &c = twcache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto key = std::make_tuple(device, 2, (long)log2n, 0L, 0L);
    auto it = c.tables.find(key);
    if (it != c.tables.end()) {
        *out = it->second;
        return ESTL_OK;
    }
    long count = rows_power_count(log2n);
    if (count < 0) return set_error(ESTL_ERR_UNSUPPORTED, "no block kernel for log2n=%d", log2n);
    float2 *d = nullptr;
    hipStream_t ts = table_stream(device);
    HIP_TRY(hipMalloc(&d, sizeof(float2) * (count > 0 ? count : 1)));
    long off = 0;
    for (int q = 1; q < rows_pass_count(log2n); ++q) {
        long A;
        int B;
        rows_pass_params(log2n, q, &A, &B);
        HIP_TRY(launch_twiddle(d + off, A, 1, A, 1, A * B, ts));
        off += A;
    }
    HIP_TRY(hipStreamSynchronize(ts));
    c.tables[key] = d;
    *out = d;
    return ESTL_OK;
}

// Table of BlockGeom<log2n>: W_{N_q}^{g s} at [(s-1) A_q + g] per pass q >= 1.
int block_table(int device, int log2n, const float2 **out) {
    auto &c = twcache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto key = std::make_tuple(device, 0, (long)log2n, 0L, 0L);
    auto it = c.tables.find(key);
    if (it != c.tables.end()) {
        *out = it->second;
        return ESTL_OK;
    }
    long count = rows_twiddle_count(log2n);
    if (count < 0) return set_error(ESTL_ERR_UNSUPPORTED, "no block kernel for log2n=%d", log2n);
    float2 *d = nullptr;
    hipStream_t ts = table_stream(device);
    HIP_TRY(hipMalloc(&d, sizeof(float2) * (count > 0 ? count : 1)));
    long off = 0;
    for (int q = 1; q < rows_pass_count(log2n); ++q) {
        long A;
        int B;
        rows_pass_params(log2n, q, &A, &B);
        long cnt = (long)(B - 1) * A;
        HIP_TRY(launch_twiddle(d + off, cnt, 0, A, 0, A * B, ts));
        off += cnt;
    }
    HIP_TRY(hipStreamSynchronize(ts));
    c.tables[key] = d;
    *out = d;
    return ESTL_OK;
}

// Linear table: entry e = exp(-2 pi i (e * mult mod nq) / nq), e < count.
int linear_table(int device, long count, long mult, long nq, const float2 **out) {
    auto &c = twcache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto key = std::make_tuple(device, 1, count, mult, nq);
    auto it = c.tables.find(key);
    if (it != c.tables.end()) {
        *out = it->second;
        return ESTL_OK;
    }
    float2 *d = nullptr;
    hipStream_t ts = table_stream(device);
    HIP_TRY(hipMalloc(&d, sizeof(float2) * count));
    HIP_TRY(launch_twiddle(d, count, 1, 1, mult, nq, ts));
    HIP_TRY(hipStreamSynchronize(ts));
    c.tables[key] = d;
    *out = d;
    return ESTL_OK;
}

static int ilog2(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

// Split log2n into passes of at most 9 bits, as evenly as possible, smaller radices
// first (a 512-point radix still leaves room for a 16-column tile in 64 KiB of LDS;
// measured on MI355X, the first pass -- the one without pre-twiddles -- is the
// cheapest place for the smaller radix).
std::vector<int> default_pass_split(int log2n) {
    int P = (log2n + 8) / 9;
    if (P < 2) P = 2;
    std::vector<int> bits(P, log2n / P);
    for (int i = 0; i < log2n % P; ++i) bits[P - 1 - i]++;
    return bits;
}

// Columns per workgroup for a pass of radix 2^log2b over `cols` columns in total:
// as wide as a 2^13-value tile (64 KiB of LDS) allows (wider tiles mean longer
// contiguous runs per load and store), capped at 64, narrowed while that leaves fewer
// than 512 tiles to spread over the CUs (small problems are latency-bound).
int default_pass_tile(int log2b, int log2cpr, long total_cols) {
    int lt = 13 - log2b;
    if (lt > kMaxPassCols) lt = kMaxPassCols;
    if (lt > log2cpr) lt = log2cpr;
    if (lt < 0) lt = 0;
    while (lt > 2 && (total_cols >> lt) < 512) --lt;
    return lt;
}

int build_plan(Plan *p, long n, long batch, int dir, int precision, unsigned flags, const std::vector<int> *split,
               const std::vector<int> *tiles) {
    if (n < 1 || (n & (n - 1)))
        return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two", n);
    if (batch < 1) return set_error(ESTL_ERR_INVALID, "batch=%ld must be >= 1", batch);
    if (dir != ESTL_FORWARD && dir != ESTL_BACKWARD)
        return set_error(ESTL_ERR_INVALID, "direction=%d must be -1 or +1", dir);
    if (precision != ESTL_PRECISION_FP32 && precision != ESTL_PRECISION_FP16)
        return set_error(ESTL_ERR_UNSUPPORTED, "precision %d not supported", precision);
    const bool fp16 = precision == ESTL_PRECISION_FP16;
    const bool planar = (flags & ESTL_FLAG_PLANAR) != 0;
    if (planar && fp16) return set_error(ESTL_ERR_UNSUPPORTED, "the tile-planarized layout is FP32 only");
    p->device = current_device();
    p->n = n;
    p->batch = batch;
    p->dir = dir;
    p->precision = precision;
    p->flags = flags;
    p->log2n = ilog2(n);
    p->scale = (flags & ESTL_FLAG_NORMALIZE) ? 1.0f / (float)n : 1.0f;
    if (fp16) {
        // FP16 storage: every pass stores 1/sqrt(radix) of its result (the rows kernel
        // 1/sqrt(n)); the output scale carries sqrt(n) and any normalisation
        p->out_factor = std::sqrt((double)n) * ((flags & ESTL_FLAG_NORMALIZE) ? 1.0 / (double)n : 1.0);
        p->scale = (float)(1.0 / std::sqrt((double)n));
    }
    if (n == 1) {
        p->kind = Plan::COPY;
        return ESTL_OK;
    }
    if (p->log2n <= kMaxRowsLog2 && !(split && !split->empty())) {
        p->kind = Plan::ROWS;
        // the rows kernels' twiddle source (kRowsTw): powers
        return power_table(p->device, p->log2n, &p->tw_rows);
    }
    p->kind = Plan::PASSES;
    std::vector<int> bits = (split && !split->empty()) ? *split : default_pass_split(p->log2n);
    int total = 0;
    const int max_b = fp16 ? max_pass_log2_fp16() : kMaxPassLog2;
    for (int b : bits) {
        const int min_b = fp16 ? 4 : 1;
        if (b < min_b || b > max_b)
            return set_error(ESTL_ERR_UNSUPPORTED, "pass radix 2^%d outside [2^%d, 2^%d]", b, min_b, max_b);
        total += b;
    }
    if (total != p->log2n) return set_error(ESTL_ERR_INVALID, "pass split does not sum to log2n=%d", p->log2n);
    if (p->log2n > 30) return set_error(ESTL_ERR_UNSUPPORTED, "n=2^%d exceeds 2^30", p->log2n);
    long A = 1;
    for (size_t q = 0; q < bits.size(); ++q) {
        PassStep st{};
        st.log2b = bits[q];
        const long B = 1L << bits[q];
        st.args.batch = batch;
        st.args.A = A;
        st.args.log2n = p->log2n;
        st.args.log2S = ilog2(n / (A * B));
        st.args.scale = fp16 ? (float)(1.0 / std::sqrt((double)B)) : (q + 1 == bits.size()) ? p->scale : 1.0f;
        st.log2t = default_pass_tile(bits[q], p->log2n - bits[q], batch * (n / B));
        st.args.xcd_remap = 0;
        if (tiles && tiles->size() == bits.size()) {
            const int lt = (*tiles)[q];
            if (lt < 0 || lt > kMaxPassCols || lt + bits[q] > kMaxPassTileLog2 || lt > p->log2n - bits[q])
                return set_error(ESTL_ERR_UNSUPPORTED, "tile 2^%d invalid for pass radix 2^%d", lt, bits[q]);
            st.log2t = lt;
            // tiles narrower than a 128-byte line share lines with their neighbours
            st.args.xcd_remap = (lt < 4) ? 1 : 0;
        }
        int rc = power_table(p->device, bits[q], &st.args.tw_local);
        if (rc) return rc;
        if (A > 1) {
            const long nq = A * B;
            const int lnq = ilog2(nq);
            st.args.lobits = (lnq + 1) / 2;
            rc = linear_table(p->device, 1L << st.args.lobits, 1, nq, &st.args.tw_lo);
            if (rc) return rc;
            rc = linear_table(p->device, nq >> st.args.lobits, 1L << st.args.lobits, nq, &st.args.tw_hi);
            if (rc) return rc;
        }
        p->steps.push_back(st);
        A *= B;
    }
    if (planar) {  // the first pass reads, the last writes the tile-planarized layout
        p->steps.front().args.planar_in = 1;
        p->steps.back().args.planar_out = 1;
    }
    p->scratch_bytes = (fp16 ? 4 : sizeof(float2)) * (size_t)n * (size_t)batch;
    return ESTL_OK;
}

int ensure_scratch(Plan *p, int which) {
    if (p->scratch[which]) return ESTL_OK;
    HIP_TRY(hipMalloc(&p->scratch[which], p->scratch_bytes));
    return ESTL_OK;
}

int stream_scratch(Plan *p, hipStream_t st, bool two, float2 **ws0, float2 **ws1) {
    std::lock_guard<std::mutex> lk(p->by_stream->mu);
    auto &pr = p->by_stream->bufs[st];
    if (!pr.first) HIP_TRY(hipMalloc(&pr.first, p->scratch_bytes));
    if (two && !pr.second) HIP_TRY(hipMalloc(&pr.second, p->scratch_bytes));
    *ws0 = pr.first;
    if (ws1) *ws1 = pr.second;
    return ESTL_OK;
}

int execute_plan(Plan *p, const float2 *in, float2 *out, hipStream_t st) {
    return execute_plan_ws(p, in, out, nullptr, nullptr, st);
}

// Same, with caller-provided workspace (each plan->scratch_bytes; NULL = the plan's own,
// allocated on first use).  ws1 is needed only for in-place odd-length chains.
int execute_plan_ws(Plan *p, const float2 *in, float2 *out, float2 *ws0, float2 *ws1, hipStream_t st) {
    if (!in || !out) return set_error(ESTL_ERR_INVALID, "null buffer");
    if (p->precision != ESTL_PRECISION_FP32)
        return set_error(ESTL_ERR_INVALID, "an FP16-storage plan runs through estl_execute_fp16");
    switch (p->kind) {
        case Plan::COPY: {
            const size_t bytes = (p->flags & ESTL_FLAG_PLANAR) ? planar_bytes(p->batch) : sizeof(float2) * p->batch;
            if (p->scale == 1.0f && in != out) HIP_TRY(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, st));
            return ESTL_OK;
        }
        case Plan::ROWS:
            if (p->flags & ESTL_FLAG_PLANAR)
                HIP_TRY(launch_rows_planar(p->log2n, p->dir, in, out, p->batch, p->tw_rows, p->scale, st));
            else
                HIP_TRY(launch_rows(p->log2n, p->dir, in, out, p->batch, p->tw_rows, p->scale, st));
            return ESTL_OK;
        case Plan::PASSES: {
            const int P = (int)p->steps.size();
            int rc = ESTL_OK;
            const bool own = !ws0;  // no caller workspace: this stream's scratch
            if (own) {
                rc = stream_scratch(p, st, false, &ws0, nullptr);
                if (rc) return rc;
            }
            // destinations backwards from out, alternating with scratch 0; the first pass
            // of a longer chain must not write its own input, so an in-place odd-length
            // chain starts in scratch 1 (a single pass may run in place: every tile
            // writes exactly the positions it read)
            std::vector<float2 *> dst(P);
            for (int q = P - 1, k = 0; q >= 0; --q, ++k) dst[q] = (k & 1) ? ws0 : out;
            if (P > 1 && dst[0] == in) {
                if (!ws1) {
                    if (!own) return set_error(ESTL_ERR_INVALID, "an in-place odd-length chain needs a second workspace");
                    float2 *w0 = nullptr;
                    rc = stream_scratch(p, st, true, &w0, &ws1);
                    if (rc) return rc;
                }
                dst[0] = ws1;
            }
            const float2 *src = in;
            for (int q = 0; q < P; ++q) {
                HIP_TRY(launch_pass(p->steps[q].log2b, p->steps[q].log2t, p->dir, src, dst[q], p->steps[q].args, st));
                src = dst[q];
            }
            return ESTL_OK;
        }
    }
    return set_error(ESTL_ERR_INVALID, "bad plan");
}

void free_stream_scratch(Plan *p) {
    if (!p->by_stream) return;
    std::lock_guard<std::mutex> lk(p->by_stream->mu);
    for (auto &kv : p->by_stream->bufs) {
        if (kv.second.first) (void)hipFree(kv.second.first);
        if (kv.second.second) (void)hipFree(kv.second.second);
    }
    p->by_stream->bufs.clear();
}

void destroy_plan(Plan *p) {
    if (!p) return;
    free_stream_scratch(p);
    for (auto &s : p->scratch)
        if (s) (void)hipFree(s);
    if (p->h_in_dev) (void)hipFree(p->h_in_dev);
    if (p->h_out_dev) (void)hipFree(p->h_out_dev);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

std::string describe_plan(const Plan *p) {
    char buf[256];
    std::string s = "{";
    snprintf(buf, sizeof buf, "\"n\": %ld, \"batch\": %ld, \"direction\": %d, \"precision\": \"%s\", ", p->n,
             p->batch, p->dir, p->precision == ESTL_PRECISION_FP16 ? "fp16" : "fp32");
    s += buf;
    const bool h = p->precision == ESTL_PRECISION_FP16;
    if (p->kind == Plan::COPY) {
        s += "\"kind\": \"copy\", \"passes\": []}";
        return s;
    }
    if (p->kind == Plan::ROWS) {
        s += h ? "\"kind\": \"rows\", \"kernel\": \"k_fft_rows_h\", \"radices\": ["
               : "\"kind\": \"rows\", \"kernel\": \"k_fft_rows\", \"radices\": [";
        for (int q = 0; q < rows_pass_count(p->log2n); ++q) {
            long A;
            int B;
            rows_pass_params(p->log2n, q, &A, &B);
            snprintf(buf, sizeof buf, "%s%d", q ? ", " : "", B);
            s += buf;
        }
        s += "], \"launches\": 1}";
        return s;
    }
    s += h ? "\"kind\": \"passes\", \"kernel\": \"k_fft_pass_h\", \"passes\": ["
           : "\"kind\": \"passes\", \"kernel\": \"k_fft_pass\", \"passes\": [";
    for (size_t q = 0; q < p->steps.size(); ++q) {
        const auto &st = p->steps[q];
        snprintf(buf, sizeof buf, "%s{\"radix\": %ld, \"A\": %ld, \"S\": %ld, \"tile\": %d}", q ? ", " : "",
                 1L << st.log2b, st.args.A, 1L << st.args.log2S, 1 << st.log2t);
        s += buf;
    }
    snprintf(buf, sizeof buf, "], \"launches\": %zu}", p->steps.size());
    s += buf;
    return s;
}

int plan_launches(const Plan *p) {
    switch (p->kind) {
        case Plan::COPY: return 0;
        case Plan::ROWS: return 1;
        default: return (int)p->steps.size();
    }
}

}  // namespace estl

using namespace estl;

extern "C" {

const char *estl_last_error(void) { return last_error(); }

int estl_version(void) { return 100; }

int estl_init(int device) {
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return set_error(ESTL_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    // the no-allocation entry point's scratch address, set up once per device here
    float2 *noalloc = nullptr;
    return noalloc_scratch(&noalloc);
}

int estl_plan_create(estl_plan_t *plan, long n, long batch, int direction, int precision, unsigned flags) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan pointer is null");
    *plan = nullptr;
    Plan *p = new Plan();
    int rc;
    if ((flags & ESTL_FLAG_AUTOTUNE) && n >= 2 && !(n & (n - 1)) &&
        (precision == ESTL_PRECISION_FP32 || precision == ESTL_PRECISION_FP16) &&
        (direction == ESTL_FORWARD || direction == ESTL_BACKWARD) && batch >= 1) {
        bool rows = false;
        std::vector<int> split, tiles;
        rc = tune_plan(n, batch, direction, precision, &rows, &split, &tiles, nullptr);
        if (!rc) rc = build_plan(p, n, batch, direction, precision, flags, rows ? nullptr : &split, rows ? nullptr : &tiles);
    } else {
        rc = build_plan(p, n, batch, direction, precision, flags, nullptr);
    }
    if (rc) {
        destroy_plan(p);
        return rc;
    }
    *plan = reinterpret_cast<estl_plan_t>(p);
    return ESTL_OK;
}

int estl_plan_create_split(estl_plan_t *plan, long n, long batch, int direction, int precision, unsigned flags,
                           int npasses, const int *log2radices, const int *log2tiles) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan pointer is null");
    *plan = nullptr;
    if (npasses < 0 || npasses > 16 || (npasses > 0 && !log2radices))
        return set_error(ESTL_ERR_INVALID, "bad pass split");
    std::vector<int> split(log2radices, log2radices + npasses);
    std::vector<int> tiles;
    if (log2tiles && npasses > 0) tiles.assign(log2tiles, log2tiles + npasses);
    Plan *p = new Plan();
    int rc = build_plan(p, n, batch, direction, precision, flags, &split, &tiles);
    if (rc) {
        destroy_plan(p);
        return rc;
    }
    *plan = reinterpret_cast<estl_plan_t>(p);
    return ESTL_OK;
}

int estl_plan_destroy(estl_plan_t plan) {
    destroy_plan(reinterpret_cast<Plan *>(plan));
    return ESTL_OK;
}

int estl_execute_device(estl_plan_t plan, const void *d_in, void *d_out, void *stream) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return execute_plan(reinterpret_cast<Plan *>(plan), static_cast<const float2 *>(d_in),
                        static_cast<float2 *>(d_out), static_cast<hipStream_t>(stream));
}

int estl_execute_host(estl_plan_t plan, const void *h_in, void *h_out) {
    if (!plan || !h_in || !h_out) return set_error(ESTL_ERR_INVALID, "null argument");
    Plan *p = reinterpret_cast<Plan *>(plan);
    const long count = p->n * p->batch;
    const size_t bytes = (p->flags & ESTL_FLAG_PLANAR) ? planar_bytes(count) : sizeof(float2) * (size_t)count;
    if (!p->stream) HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    if (!p->h_in_dev) HIP_TRY(hipMalloc(&p->h_in_dev, bytes));
    if (!p->h_out_dev) HIP_TRY(hipMalloc(&p->h_out_dev, bytes));
    HIP_TRY(hipMemcpyAsync(p->h_in_dev, h_in, bytes, hipMemcpyHostToDevice, p->stream));
    int rc = execute_plan(p, p->h_in_dev, p->h_out_dev, p->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, p->h_out_dev, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return ESTL_OK;
}

long long estl_max_batch(long n, int precision, long long budget_bytes) {
    if (n < 1 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two", n);
    if (budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return set_error(ESTL_ERR_HIP, "hipMemGetInfo failed");
        budget_bytes = (long long)fr;
    }
    const long long elem = (precision == ESTL_PRECISION_FP16) ? 4 : 8;
    // input + output (+ one scratch row set for multi-pass sizes)
    const long long per_row = elem * (long long)n * ((ilog2(n) <= kMaxRowsLog2) ? 2 : 3);
    return budget_bytes / per_row;
}

long long estl_plan_workspace_bytes(estl_plan_t plan) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return (long long)reinterpret_cast<Plan *>(plan)->scratch_bytes;
}

int estl_plan_describe(estl_plan_t plan, char *buf, int len) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    std::string s = describe_plan(reinterpret_cast<Plan *>(plan));
    if (buf && len > 0) {
        strncpy(buf, s.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)s.size();
}

int estl_plan_launches(estl_plan_t plan) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return plan_launches(reinterpret_cast<Plan *>(plan));
}

int estl_stream_copy(const void *d_src, void *d_dst, long long bytes, int nontemporal, void *stream) {
    if (!d_src || !d_dst || bytes < 0) return set_error(ESTL_ERR_INVALID, "bad copy arguments");
    if ((bytes & 15) || ((reinterpret_cast<uintptr_t>(d_src) | reinterpret_cast<uintptr_t>(d_dst)) & 15))
        return set_error(ESTL_ERR_INVALID, "the copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    HIP_TRY(launch_stream_copy(d_src, d_dst, (long)(bytes / 16), nontemporal != 0, static_cast<hipStream_t>(stream)));
    return ESTL_OK;
}

int estl_fill_uniform(void *d_dst, long count, uint64_t seed, uint64_t offset, void *stream) {
    if (!d_dst && count > 0) return set_error(ESTL_ERR_INVALID, "null buffer");
    HIP_TRY(launch_fill_uniform(static_cast<float *>(d_dst), count, seed, offset, static_cast<hipStream_t>(stream)));
    return ESTL_OK;
}

}  // extern "C"
