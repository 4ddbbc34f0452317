// tune.cpp -- autotuner over the plan space.
//
// The reference hard-codes one plan per size: radix-16 passes with the remainder last
// (FFT_OpenCL_Contiguous.h:55-60), each pass a separate generated kernel.  Here the plan
// space is (a) the single-launch rows kernel when n <= 2^13 and (b) every split of log2 n
// into 2..4 global passes of radix 2^4..2^13 near the balanced split, each with its
// default tile width, then the tile widths of the three fastest splits and of the
// fastest split with the fewest passes refined pass by pass.  The tuner times each
// candidate on the device (HIP events, min of a few runs on zero-filled buffers of the
// real size) and keeps the fastest per (device, log2 n, batch, direction), preferring
// fewer passes within 2 %.  Choices persist, one line per key, in the text file named
// by ESTL_TUNE_CACHE when that variable is set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/estl_hip.h"
#include "plan.h"

namespace estl {

namespace {

struct Choice {
    bool rows = false;
    std::vector<int> split, tiles;
    double us = 0;
};

using Key = std::tuple<std::string, int, long, int>;  // arch, log2n, batch, dir

std::mutex g_mu;
std::map<Key, Choice> g_cache;
bool g_loaded = false;

std::string device_arch(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

const char *cache_path() {
    const char *p = getenv("ESTL_TUNE_CACHE");
    return (p && *p) ? p : nullptr;
}

// line: arch log2n batch dir rows|passes P s_1..s_P t_1..t_P us
void load_cache_locked() {
    if (g_loaded) return;
    g_loaded = true;
    const char *path = cache_path();
    if (!path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char arch[64], kind[16];
        int log2n, dir, P, off = 0;
        long batch;
        if (sscanf(line, "%63s %d %ld %d %15s %d%n", arch, &log2n, &batch, &dir, kind, &P, &off) != 6) continue;
        if (P < 0 || P > 16) continue;
        Choice c;
        c.rows = strcmp(kind, "rows") == 0;
        const char *q = line + off;
        bool ok = true;
        for (int k = 0; k < 2 * P && ok; ++k) {
            int v, adv;
            if (sscanf(q, "%d%n", &v, &adv) != 1) ok = false;
            else {
                (k < P ? c.split : c.tiles).push_back(v);
                q += adv;
            }
        }
        if (!ok || sscanf(q, "%lf", &c.us) != 1) continue;
        g_cache[Key(arch, log2n, batch, dir)] = c;
    }
    fclose(f);
}

void save_cache_locked() {
    const char *path = cache_path();
    if (!path) return;
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    for (const auto &kv : g_cache) {
        const Choice &c = kv.second;
        fprintf(f, "%s %d %ld %d %s %zu", std::get<0>(kv.first).c_str(), std::get<1>(kv.first), std::get<2>(kv.first),
                std::get<3>(kv.first), c.rows ? "rows" : "passes", c.split.size());
        for (int v : c.split) fprintf(f, " %d", v);
        for (int v : c.tiles) fprintf(f, " %d", v);
        fprintf(f, " %.3f\n", c.us);
    }
    fclose(f);
    rename(tmp.c_str(), path);
}

// all compositions of L into P parts within [lo, hi]
void compositions(int L, int P, int lo, int hi, std::vector<int> &cur, std::vector<std::vector<int>> &out) {
    if (P == 0) {
        if (L == 0) out.push_back(cur);
        return;
    }
    for (int v = lo; v <= hi && v <= L; ++v) {
        cur.push_back(v);
        compositions(L - v, P - 1, lo, hi, cur, out);
        cur.pop_back();
    }
}

std::vector<Choice> candidates(int log2n, int max_b) {
    std::vector<Choice> out;
    if (log2n <= kMaxRowsLog2) {
        Choice c;
        c.rows = true;
        out.push_back(c);
    }
    const int p0 = std::max(2, (log2n + 8) / 9);
    for (int P = std::max(2, p0 - 1); P <= std::min(4, p0 + 1); ++P) {
        if (P * 4 > log2n || P * max_b < log2n) continue;
        const int lo = std::max(4, log2n / P - 1), hi = std::min(max_b, (log2n + P - 1) / P + 1);
        std::vector<std::vector<int>> comps;
        std::vector<int> cur;
        compositions(log2n, P, lo, hi, cur, comps);
        for (auto &s : comps) {
            Choice c;
            c.split = s;
            out.push_back(c);
        }
    }
    return out;
}

}  // namespace

int default_pass_tile(int log2b, int log2cpr, long total_cols);

int execute_plan_fp16(Plan *p, const void *in, const float *in_scale, void *out, float *out_scale, hipStream_t st);

// Time one candidate: median-free minimum of `reps` runs after a warm-up.  FP16
// candidates run through the FP16 entry point (`scales`: two arrays of batch floats).
static int time_candidate(const Choice &c, long n, long batch, int dir, int precision, float2 *in, float2 *out,
                          float *scales, hipStream_t st, hipEvent_t e0, hipEvent_t e1, double *us) {
    Plan p;
    int rc = build_plan(&p, n, batch, dir, precision, 0, c.rows ? nullptr : &c.split,
                        c.tiles.empty() ? nullptr : &c.tiles);
    if (rc) return rc;
    double best = 1e30;
    const int reps = 4;
    for (int r = 0; r < reps + 1; ++r) {
        if (hipEventRecord(e0, st) != hipSuccess) return set_error(ESTL_ERR_HIP, "event record failed");
        rc = precision == ESTL_PRECISION_FP16 ? execute_plan_fp16(&p, in, scales, out, scales + batch, st)
                                              : execute_plan(&p, in, out, st);
        if (rc) break;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return set_error(ESTL_ERR_HIP, "event sync failed");
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return set_error(ESTL_ERR_HIP, "elapsed time failed");
        if (r > 0) best = std::min(best, 1e3 * (double)ms);
    }
    for (auto &s : p.scratch)
        if (s) (void)hipFree(s);
    p.scratch[0] = p.scratch[1] = nullptr;
    free_stream_scratch(&p);
    *us = best;
    return rc;
}

// Tune (or look up) the plan for (n, batch, dir) on the current device.  On success
// fills split/tiles (empty split = rows kernel) and, if report is given, a JSON report
// of every candidate's time.
int tune_plan(long n, long batch, int dir, int precision, bool *rows, std::vector<int> *split,
              std::vector<int> *tiles, std::string *report) {
    if (n < 2 || (n & (n - 1))) return set_error(ESTL_ERR_INVALID, "n=%ld is not a power of two >= 2", n);
    int log2n = 0;
    while ((1L << log2n) < n) ++log2n;
    const int device = current_device();
    if (precision != ESTL_PRECISION_FP32 && precision != ESTL_PRECISION_FP16)
        return set_error(ESTL_ERR_UNSUPPORTED, "precision %d not supported", precision);
    const bool fp16 = precision == ESTL_PRECISION_FP16;
    // FP16 plans are cached under the architecture name plus "+fp16"
    const Key key(device_arch(device) + (fp16 ? "+fp16" : ""), log2n, batch, dir);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_cache_locked();
        auto it = g_cache.find(key);
        if (it != g_cache.end() && !report) {
            *rows = it->second.rows;
            *split = it->second.split;
            *tiles = it->second.tiles;
            return ESTL_OK;
        }
    }
    const size_t bytes = sizeof(float2) * (size_t)n * (size_t)batch;
    float2 *in = nullptr, *out = nullptr;
    float *scales = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ESTL_OK;
    auto cleanup = [&] {
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (scales) (void)hipFree(scales);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    if (hipMalloc(&in, bytes) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess ||
        hipMalloc(&scales, 2 * sizeof(float) * (size_t)batch) != hipSuccess) {
        cleanup();
        return set_error(ESTL_ERR_NOMEM, "tuner could not allocate 2 x %zu bytes", bytes);
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(in, 0, bytes, st) != hipSuccess ||
        hipMemsetAsync(scales, 0, 2 * sizeof(float) * (size_t)batch, st) != hipSuccess) {
        cleanup();
        return set_error(ESTL_ERR_HIP, "tuner set-up failed");
    }
    std::vector<Choice> cands = candidates(log2n, fp16 ? max_pass_log2_fp16() : kMaxPassLog2);
    Choice best;
    best.us = 1e30;
    std::vector<Choice> timed;  // pass splits that ran, fastest first after the first stage
    std::string rep = "{\"n\": " + std::to_string(n) + ", \"batch\": " + std::to_string(batch) +
                      ", \"direction\": " + std::to_string(dir) + ", \"candidates\": [";
    bool first = true;
    for (auto &c : cands) {
        double us = 0;
        rc = time_candidate(c, n, batch, dir, precision, in, out, scales, st, e0, e1, &us);
        if (rc) continue;  // a split without kernels for this size: skip it
        c.us = us;
        if (!c.rows) {
            // record the tile widths the plan used
            c.tiles.clear();
            for (int b : c.split) c.tiles.push_back(default_pass_tile(b, log2n - b, batch * (n >> b)));
        }
        char buf[256];
        std::string s = c.rows ? "\"rows\"" : "[";
        if (!c.rows) {
            for (size_t k = 0; k < c.split.size(); ++k) s += (k ? ", " : "") + std::to_string(c.split[k]);
            s += "]";
        }
        snprintf(buf, sizeof buf, "%s{\"plan\": %s, \"us\": %.2f}", first ? "" : ", ", s.c_str(), us);
        rep += buf;
        first = false;
        if (us < best.us) best = c;
        if (!c.rows) timed.push_back(c);
    }
    std::sort(timed.begin(), timed.end(), [](const Choice &a, const Choice &b) { return a.us < b.us; });
    // second stage: tile widths, one pass at a time (coordinate descent over 2^-2 ..
    // 2^+1 times the default width; narrow tiles get the XCD remap), for the three
    // fastest splits and the fastest split of the fewest passes, so a split whose
    // default tiles are poor still competes; a plan of fewer passes (fewer HBM round
    // trips) is kept when it is within 2 % of the fastest
    if (best.us < 1e30) {
        std::vector<Choice> seeds;
        for (const Choice &c : timed) {
            if ((int)seeds.size() < 3) seeds.push_back(c);
        }
        size_t fewest = 0;
        for (size_t i = 1; i < timed.size(); ++i)
            if (timed[i].split.size() < timed[fewest].split.size()) fewest = i;
        if (!timed.empty() && fewest >= seeds.size()) seeds.push_back(timed[fewest]);
        std::vector<Choice> refined;
        for (Choice cur : seeds) {
            for (size_t q = 0; q < cur.split.size(); ++q) {
                const int b = cur.split[q], d = cur.tiles[q];
                for (int lt : {d - 2, d - 1, d + 1}) {
                    if (lt < 0 || lt > kMaxPassCols || lt + b > kMaxPassTileLog2 || lt > log2n - b) continue;
                    Choice c = cur;
                    c.tiles[q] = lt;
                    double us = 0;
                    if (time_candidate(c, n, batch, dir, precision, in, out, scales, st, e0, e1, &us)) continue;
                    char buf[256];
                    std::string s = "[";
                    for (size_t k = 0; k < c.split.size(); ++k) s += (k ? ", " : "") + std::to_string(c.split[k]);
                    s += "], \"tiles\": [";
                    for (size_t k = 0; k < c.tiles.size(); ++k) s += (k ? ", " : "") + std::to_string(c.tiles[k]);
                    s += "]";
                    snprintf(buf, sizeof buf, ", {\"plan\": %s, \"us\": %.2f}", s.c_str(), us);
                    rep += buf;
                    if (us < cur.us) {
                        c.us = us;
                        cur = c;
                    }
                }
            }
            refined.push_back(cur);
            if (cur.us < best.us) best = cur;
        }
        for (const Choice &c : refined)
            if (!best.rows && c.split.size() < best.split.size() && c.us <= 1.02 * best.us) best = c;
    }
    cleanup();
    rep += "]}";
    if (best.us >= 1e30) return set_error(ESTL_ERR_UNSUPPORTED, "no candidate plan ran for n=%ld", n);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = best;
        save_cache_locked();
    }
    *rows = best.rows;
    *split = best.split;
    *tiles = best.tiles;
    if (report) *report = rep;
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" int estl_tune(long n, long batch, int direction, char *report, int len) {
    return estl_tune_precision(n, batch, direction, ESTL_PRECISION_FP32, report, len);
}

extern "C" int estl_tune_precision(long n, long batch, int direction, int precision, char *report, int len) {
    bool rows = false;
    std::vector<int> split, tiles;
    std::string rep;
    const int rc = tune_plan(n, batch, direction, precision, &rows, &split, &tiles, &rep);
    if (rc) return rc;
    if (report && len > 0) {
        strncpy(report, rep.c_str(), (size_t)len - 1);
        report[len - 1] = 0;
    }
    return (int)rep.size();
}
