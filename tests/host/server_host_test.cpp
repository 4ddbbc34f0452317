// server_host_test.cpp -- the request server (estl_amd/csrc/server.cpp) on the CPU:
// built unchanged over the fake HIP runtime (fake_hip/) and the CPU stand-ins of the
// launch functions (fake_launch.cpp), driven through its C ABI (estl_server_*), and run
// under ThreadSanitizer or AddressSanitizer + UBSan by tests/test_server_host.py.
//
// The reference has no serving layer; its error path prints and carries on
// (FFT_OpenCL_Contiguous.h:120-122, uOpenCL.h:21-55).  Checked here:
//   mix       a seeded stream of FFT and convolution requests (FP32 and FP16 storage,
//             inline and registered filters, single rows, request groups) submitted from
//             two threads to a 4-stream server: every result equals the same request
//             run alone, bit for bit, and FP32 results equal a direct FP64 reference;
//   denial    a request larger than the budget fails with ESTL_ERR_NOMEM; a malformed
//             one (n not a power of two, unknown flags) with ESTL_ERR_INVALID; the
//             server keeps serving;
//   launchfail the k-th launch of the mixed stream fails when issued, for several k:
//             the requests it belonged to (a whole group) fail with ESTL_ERR_HIP, every
//             other request equals its solo result, bit for bit;
//   asyncfail the k-th launch fails on the device: the same, through the event path;
//   stale     a request whose second launch fails while its first is still queued:
//             its memory must not go to the next request before that first launch has
//             run (it would write into the next request's output);
//   copyout   estl_server_copy_output on a stalled stream, then release, then a new
//             request: the copy must still see the released output;
//   shutdown  destroy with requests queued and running.
// Usage: server_host_test <case> [seed]; prints PASS/FAIL lines, exits 0 when all pass.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/estl_hip.h"

extern "C" {
void fake_set_fail_launch(long k);
void fake_set_fail_async(long k);
void fake_reset_launches();
long fake_launches();
void fake_set_stall_us(int us);
void fake_stall_next_task_us(int us);
}

static int g_failures = 0;
static long g_groups = 0;  // request groups the mixed runs formed
#define CHECK(cond, ...)                     \
    do {                                     \
        if (!(cond)) {                       \
            printf("FAIL ");                 \
            printf(__VA_ARGS__);             \
            printf("\n");                    \
            ++g_failures;                    \
        }                                    \
    } while (0)

enum Kind { FFT = 0, CONV = 1, CONV_REG = 2, CONV16 = 3, CONV16_REG = 4 };

struct Spec {
    int kind;
    long n, rows, k;
    int dir;
    unsigned flags;
    int filter;  // index into the registered filters (CONV_REG / CONV16_REG)
    std::vector<float> x, h, xscale;
    std::vector<uint16_t> x16;
    size_t out_bytes() const {
        if (kind == FFT) return 8 * (size_t)n * rows;
        if (kind == CONV || kind == CONV_REG) return 4 * (size_t)n * rows;
        return 2 * (size_t)n * rows;  // binary16 rows (scales compared separately)
    }
};

struct FilterSpec {
    long n, k;
    unsigned flags;
    bool fp16;
    std::vector<float> h;
};

struct Result {
    int rc = 1;
    std::vector<unsigned char> out;
    std::vector<float> scales;
};

static std::vector<FilterSpec> make_filters(std::mt19937_64 &g) {
    std::vector<FilterSpec> f;
    const long ns[] = {64, 1024, 1L << 15};
    for (long n : ns)
        for (int fp16 = 0; fp16 < 2; ++fp16) {
            FilterSpec s;
            s.n = n;
            s.k = 1 + (long)(g() % 40);
            s.flags = 0;
            s.fp16 = fp16;
            std::uniform_real_distribution<float> u(-1, 1);
            for (long i = 0; i < s.k; ++i) s.h.push_back(u(g));
            f.push_back(s);
        }
    return f;
}

static std::vector<Spec> make_stream(std::mt19937_64 &g, const std::vector<FilterSpec> &filters, int count) {
    std::vector<Spec> v;
    std::uniform_real_distribution<float> u(-1, 1);
    const long ns[] = {16, 64, 1024, 1L << 14, 1L << 15};
    for (int i = 0; i < count; ++i) {
        Spec s{};
        s.kind = (int)(g() % 5);
        s.n = ns[g() % 5];
        s.rows = (g() % 3 == 0) ? 1 : 1 + (long)(g() % 6);
        s.dir = (g() & 1) ? ESTL_FORWARD : ESTL_BACKWARD;
        s.flags = (g() % 4 == 0) ? ESTL_CONV_CORRELATE : 0;
        if (s.kind == CONV && g() % 8 == 0) s.flags |= ESTL_CONV_CENTER;
        if (s.kind == CONV && g() % 16 == 0) s.flags |= ESTL_CONV_NORMALIZE_KERNEL;
        if (s.kind == CONV_REG || s.kind == CONV16_REG) {
            // a registered filter of the right precision; its n is the request's
            std::vector<int> ok;
            for (size_t f = 0; f < filters.size(); ++f)
                if (filters[f].fp16 == (s.kind == CONV16_REG)) ok.push_back((int)f);
            s.filter = ok[g() % ok.size()];
            s.n = filters[s.filter].n;
        }
        if (s.kind == FFT && s.n > 4096) s.rows = 1;
        s.k = 1 + (long)(g() % 100);
        const long vals = s.n * s.rows;
        if (s.kind == FFT)
            for (long j = 0; j < 2 * vals; ++j) s.x.push_back(u(g));
        else if (s.kind == CONV || s.kind == CONV_REG)
            for (long j = 0; j < vals; ++j) s.x.push_back(u(g));
        else {
            for (long j = 0; j < vals; ++j) s.x16.push_back((uint16_t)(g() & 0x7bff));
            for (long r = 0; r < s.rows; ++r) s.xscale.push_back(0.5f + (float)(g() % 100) / 64.0f);
        }
        if (s.kind == CONV || s.kind == CONV16)
            for (long j = 0; j < s.k; ++j) s.h.push_back(u(g));
        v.push_back(std::move(s));
    }
    return v;
}

// input buffers stay alive (and unchanged) for the whole run: device memory here is host
// memory, so the request's pointers are the vectors' data
static int submit(estl_server_t srv, Spec &s, const std::vector<long long> &filter_ids, long long *id) {
    switch (s.kind) {
        case FFT:
            return estl_server_submit_fft(srv, s.n, s.rows, s.dir, s.x.data(), nullptr, id);
        case CONV:
            return estl_server_submit_conv(srv, s.n, s.rows, s.x.data(), s.h.data(), s.k, s.flags, nullptr, id);
        case CONV_REG:
            return estl_server_submit_conv_filter(srv, filter_ids[s.filter], s.rows, s.x.data(),
                                                  s.flags & ESTL_CONV_CORRELATE, nullptr, id);
        case CONV16:
            return estl_server_submit_conv_fp16(srv, s.n, s.rows, s.x16.data(), s.xscale.data(), s.h.data(), s.k,
                                                s.flags & ESTL_CONV_CORRELATE, nullptr, id);
        case CONV16_REG:
            return estl_server_submit_conv_filter_fp16(srv, filter_ids[s.filter], s.rows, s.x16.data(),
                                                       s.xscale.data(), s.flags & ESTL_CONV_CORRELATE, nullptr, id);
    }
    return ESTL_ERR_INVALID;
}

static Result collect(estl_server_t srv, const Spec &s, long long id) {
    Result res;
    int state = 0;
    void *out = nullptr;
    res.rc = estl_server_poll(srv, id, &state, &out, nullptr);
    if (res.rc == ESTL_OK && state == 2) {
        const unsigned char *p = static_cast<const unsigned char *>(out);
        res.out.assign(p, p + s.out_bytes());
        if (s.kind == CONV16 || s.kind == CONV16_REG) {
            const size_t off = (2 * (size_t)s.n * (size_t)s.rows + 255) & ~size_t(255);
            const float *sc = reinterpret_cast<const float *>(p + off);
            res.scales.assign(sc, sc + s.rows);
        }
    }
    return res;
}

static std::vector<long long> register_filters(estl_server_t srv, std::vector<FilterSpec> &filters) {
    std::vector<long long> ids;
    for (auto &f : filters) {
        long long id = 0;
        int rc = estl_server_register_filter(srv, f.n, f.h.data(), f.k, f.fp16 ? ESTL_CONV_FP16 : 0, nullptr, &id);
        CHECK(rc == ESTL_OK, "register filter n=%ld: %s", f.n, estl_last_error());
        rc = estl_server_wait(srv, id, 30.0);
        CHECK(rc == ESTL_OK, "filter %lld: %s", id, estl_last_error());
        ids.push_back(id);
    }
    return ids;
}

// every request on a one-stream server, one at a time
static std::vector<Result> run_alone(std::vector<Spec> &reqs, std::vector<FilterSpec> &filters) {
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, 64LL << 20, 1) == ESTL_OK, "create: %s", estl_last_error());
    std::vector<long long> fids = register_filters(srv, filters);
    std::vector<Result> out(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) {
        long long id = 0;
        int rc = submit(srv, reqs[i], fids, &id);
        if (rc == ESTL_OK) rc = estl_server_wait(srv, id, 30.0);
        out[i] = rc == ESTL_OK ? collect(srv, reqs[i], id) : Result{rc, {}, {}};
        if (id) estl_server_release(srv, id);
    }
    for (long long f : fids) estl_server_release(srv, f);
    estl_server_destroy(srv);
    return out;
}

// the stream submitted from two threads to a server of `streams` streams; results are
// read once everything has finished
// (fail_launch / fail_async: the launch of the stream, counted after the filters are
// registered, that fails when issued / on the device; 0 = none)
static std::vector<Result> run_mix(std::vector<Spec> &reqs, std::vector<FilterSpec> &filters, int streams,
                                   long long budget, bool release_early, long fail_launch = 0, long fail_async = 0) {
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, budget, streams) == ESTL_OK, "create: %s", estl_last_error());
    std::vector<long long> fids = register_filters(srv, filters);
    fake_reset_launches();
    fake_set_fail_launch(fail_launch);
    fake_set_fail_async(fail_async);
    std::vector<long long> ids(reqs.size(), 0);
    std::vector<int> sub_rc(reqs.size(), ESTL_OK);
    auto submitter = [&](int part) {
        for (size_t i = part; i < reqs.size(); i += 2) sub_rc[i] = submit(srv, reqs[i], fids, &ids[i]);
    };
    std::thread t0(submitter, 0), t1(submitter, 1);
    t0.join();
    t1.join();
    std::vector<Result> out(reqs.size());
    std::map<long long, size_t> index;
    for (size_t i = 0; i < reqs.size(); ++i) {
        if (sub_rc[i] != ESTL_OK) out[i].rc = sub_rc[i];
        else index[ids[i]] = i;
    }
    // completions as they come (next_done); results read at once and released when asked
    size_t seen = 0;
    while (seen < index.size()) {
        long long done[64];
        const int c = estl_server_next_done(srv, done, 64, 30.0);
        CHECK(c > 0, "next_done timed out with %zu of %zu seen", seen, index.size());
        if (c <= 0) break;
        for (int j = 0; j < c; ++j) {
            auto it = index.find(done[j]);
            if (it == index.end()) continue;
            const size_t i = it->second;
            out[i] = collect(srv, reqs[i], done[j]);
            ++seen;
            if (release_early) estl_server_release(srv, done[j]);
        }
    }
    CHECK(estl_server_drain(srv, 30.0) == ESTL_OK, "drain: %s", estl_last_error());
    fake_set_fail_launch(0);
    fake_set_fail_async(0);
    char js[1024];
    estl_server_stats(srv, js, sizeof js);
    if (const char *g = strstr(js, "\"groups\": ")) g_groups += atol(g + 10);
    if (!release_early)
        for (auto &kv : index) estl_server_release(srv, kv.first);
    for (long long f : fids) estl_server_release(srv, f);
    estl_server_destroy(srv);
    return out;
}

// FP64 references: the DFT of an FFT request, the circular convolution of an FP32 one
static double check_reference(const Spec &s, const Result &r) {
    using cd = std::complex<double>;
    double worst = 0;
    if (s.kind == FFT) {
        if (s.n > 1024) return 0;
        const float *y = reinterpret_cast<const float *>(r.out.data());
        for (long row = 0; row < s.rows; ++row) {
            double num = 0, den = 0;
            for (long f = 0; f < s.n; ++f) {
                cd acc = 0;
                for (long t = 0; t < s.n; ++t) {
                    const double a = s.dir * 2.0 * M_PI * (double)((f * t) % s.n) / (double)s.n;
                    acc += cd(s.x[2 * (row * s.n + t)], s.x[2 * (row * s.n + t) + 1]) * cd(std::cos(a), std::sin(a));
                }
                const cd got(y[2 * (row * s.n + f)], y[2 * (row * s.n + f) + 1]);
                num += std::norm(got - acc);
                den += std::norm(acc);
            }
            worst = std::max(worst, std::sqrt(num / std::max(den, 1e-300)));
        }
        return worst;
    }
    if (s.kind != CONV || (s.flags & ESTL_CONV_NORMALIZE_KERNEL) || s.n > 1024) return 0;
    const long shift = (s.flags & ESTL_CONV_CENTER) ? ((s.k + 1) / 2) % s.n : 0;
    std::vector<double> f(s.n, 0.0);
    for (long i = 0; i < s.k; ++i) f[((i - shift) % s.n + s.n) % s.n] += s.h[i];
    const float *y = reinterpret_cast<const float *>(r.out.data());
    for (long row = 0; row < s.rows; ++row) {
        double num = 0, den = 0;
        for (long t = 0; t < s.n; ++t) {
            double acc = 0;
            for (long j = 0; j < s.n; ++j) {
                const long m = (s.flags & ESTL_CONV_CORRELATE) ? (t + j) % s.n : ((t - j) % s.n + s.n) % s.n;
                acc += f[j] * s.x[row * s.n + m];
            }
            num += (y[row * s.n + t] - acc) * (y[row * s.n + t] - acc);
            den += acc * acc;
        }
        worst = std::max(worst, std::sqrt(num / std::max(den, 1e-300)));
    }
    return worst;
}

static bool same(const Result &a, const Result &b) { return a.rc == b.rc && a.out == b.out && a.scales == b.scales; }

static void case_mix(unsigned seed) {
    std::mt19937_64 g(seed);
    auto filters = make_filters(g);
    auto reqs = make_stream(g, filters, 160);
    const auto alone = run_alone(reqs, filters);
    int ok = 0;
    double worst = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        CHECK(alone[i].rc == ESTL_OK, "alone request %zu (kind %d n %ld rows %ld): rc %d", i, reqs[i].kind,
              reqs[i].n, reqs[i].rows, alone[i].rc);
        if (alone[i].rc == ESTL_OK) worst = std::max(worst, check_reference(reqs[i], alone[i]));
    }
    CHECK(worst < 1e-5, "FP64 reference: worst relative error %.3e", worst);
    for (int pass = 0; pass < 2; ++pass) {
        const auto mix = run_mix(reqs, filters, 4, 64LL << 20, pass == 1);
        for (size_t i = 0; i < reqs.size(); ++i) {
            CHECK(same(alone[i], mix[i]), "mix (release %s) request %zu kind %d n %ld rows %ld differs from alone",
                  pass ? "early" : "late", i, reqs[i].kind, reqs[i].n, reqs[i].rows);
            ok += same(alone[i], mix[i]);
        }
    }
    CHECK(g_groups > 0, "the mixed runs formed no request group");
    printf("%s mix: %d/%zu results equal alone, FP64 worst %.2e, %ld request groups\n", g_failures ? "FAIL" : "PASS",
           ok, 2 * reqs.size(), worst, g_groups);
}

static void case_denial() {
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, 1 << 20, 2) == ESTL_OK, "create");
    std::vector<float> big(2 * (1 << 17)), small(2 * 256);
    for (size_t i = 0; i < small.size(); ++i) small[i] = (float)(i % 7) - 3.0f;
    long long a = 0, b = 0, c = 0, d = 0;
    CHECK(estl_server_submit_fft(srv, 1 << 17, 1, ESTL_FORWARD, big.data(), nullptr, &a) == ESTL_OK, "submit big");
    CHECK(estl_server_submit_fft(srv, 12, 1, ESTL_FORWARD, small.data(), nullptr, &b) == ESTL_OK, "submit n=12");
    std::vector<float> h(5, 1.0f);
    CHECK(estl_server_submit_conv(srv, 256, 1, small.data(), h.data(), 5, 0x1000u, nullptr, &c) == ESTL_OK,
          "submit bad flags");
    CHECK(estl_server_submit_fft(srv, 256, 1, ESTL_FORWARD, small.data(), nullptr, &d) == ESTL_OK, "submit ok");
    const int ra = estl_server_wait(srv, a, 10), rb = estl_server_wait(srv, b, 10), rc = estl_server_wait(srv, c, 10),
              rd = estl_server_wait(srv, d, 10);
    CHECK(ra == ESTL_ERR_NOMEM, "a request over the budget: %d (%s)", ra, estl_last_error());
    CHECK(rb == ESTL_ERR_INVALID, "n = 12: %d", rb);
    CHECK(rc == ESTL_ERR_INVALID, "unknown flags: %d", rc);
    CHECK(rd == ESTL_OK, "the next request is served: %d", rd);
    char js[1024];
    estl_server_stats(srv, js, sizeof js);
    CHECK(strstr(js, "\"rejected\": 1") != nullptr, "stats count one rejection: %s", js);
    CHECK(estl_server_release(srv, a) == ESTL_OK && estl_server_release(srv, d) == ESTL_OK, "release");
    long long e = 0;
    CHECK(estl_server_submit_fft(srv, 64, 1, 0, small.data(), nullptr, &e) == ESTL_OK, "submit bad direction");
    CHECK(estl_server_wait(srv, e, 10) == ESTL_ERR_INVALID, "direction 0 is rejected");
    estl_server_destroy(srv);
    printf("%s denial\n", g_failures ? "FAIL" : "PASS");
}

static void case_launch_failure(unsigned seed, bool async) {
    std::mt19937_64 g(seed);
    auto filters = make_filters(g);
    auto reqs = make_stream(g, filters, 60);
    const auto alone = run_alone(reqs, filters);
    // count the launches of the failure-free mixed run, then fail several of them
    fake_set_stall_us(200);
    (void)run_mix(reqs, filters, 3, 16LL << 20, true);
    const long total = fake_launches();
    for (long k : {1L, 3L, total / 3, total / 2, total - 2}) {
        if (k < 1) continue;
        const auto mix = run_mix(reqs, filters, 3, 16LL << 20, true, async ? 0 : k, async ? k : 0);
        // how requests group depends on timing, so a run may issue fewer launches than k
        const bool hit = fake_launches() >= k;
        int failed = 0, wrong = 0;
        for (size_t i = 0; i < reqs.size(); ++i) {
            if (mix[i].rc != ESTL_OK) {
                ++failed;
                CHECK(mix[i].rc == ESTL_ERR_HIP, "launch %ld: request %zu failed with %d, not ESTL_ERR_HIP", k, i,
                      mix[i].rc);
            } else if (!same(mix[i], alone[i])) {
                ++wrong;
            }
        }
        CHECK(failed >= 1 || !hit, "launch %ld of %ld: no request failed", k, total);
        CHECK(wrong == 0, "launch %ld of %ld: %d surviving requests differ from alone", k, total, wrong);
        printf("%s %s failure of launch %ld/%ld: %d failed, %d wrong%s\n", wrong || (hit && !failed) ? "FAIL" : "PASS",
               async ? "device" : "issue", k, total, failed, wrong, hit ? "" : " (this run issued fewer launches)");
    }
    fake_set_stall_us(0);
}

// A convolution (output 4 KiB, spectrum 8 KiB, ...) whose second launch fails while its
// first (the filter spectrum, held up 300 ms) is still queued; then an FFT whose 16 KiB
// output only fits once the convolution's memory is free.  The FFT's output must not
// change after it completes.
static void case_stale() {
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, 16 << 10, 2) == ESTL_OK, "create");
    std::vector<float> x(1024), h(16), z(2 * 2048);
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> u(-1, 1);
    for (auto &v : x) v = u(g);
    for (auto &v : h) v = u(g);
    for (auto &v : z) v = u(g);
    fake_reset_launches();
    fake_stall_next_task_us(300000);
    fake_set_fail_launch(2);
    long long a = 0, b = 0;
    CHECK(estl_server_submit_conv(srv, 1024, 1, x.data(), h.data(), 16, 0, nullptr, &a) == ESTL_OK, "submit conv");
    const int ra = estl_server_wait(srv, a, 10);
    fake_set_fail_launch(0);
    CHECK(ra == ESTL_ERR_HIP, "the convolution fails with ESTL_ERR_HIP: %d", ra);
    CHECK(estl_server_submit_fft(srv, 2048, 1, ESTL_FORWARD, z.data(), nullptr, &b) == ESTL_OK, "submit fft");
    CHECK(estl_server_wait(srv, b, 10) == ESTL_OK, "fft done: %s", estl_last_error());
    Spec s{};
    s.kind = FFT;
    s.n = 2048;
    s.rows = 1;
    const Result first = collect(srv, s, b);
    std::this_thread::sleep_for(std::chrono::milliseconds(400));  // past the held-up launch
    const Result later = collect(srv, s, b);
    CHECK(first.rc == ESTL_OK && same(first, later), "the FFT's output changed after completion");
    estl_server_release(srv, a);
    estl_server_release(srv, b);
    estl_server_destroy(srv);
    printf("%s stale\n", g_failures ? "FAIL" : "PASS");
}

// copy_output queued on a stalled stream, release at once, then a request that reuses
// the block: the copy must still read the first output.
static void case_copy_output() {
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, 16 << 10, 2) == ESTL_OK, "create");
    std::vector<float> z1(2 * 2048), z2(2 * 2048), dst(2 * 2048, 0.0f);
    std::mt19937_64 g(8);
    std::uniform_real_distribution<float> u(-1, 1);
    for (auto &v : z1) v = u(g);
    for (auto &v : z2) v = u(g);
    long long a = 0, b = 0;
    CHECK(estl_server_submit_fft(srv, 2048, 1, ESTL_FORWARD, z1.data(), nullptr, &a) == ESTL_OK, "submit");
    CHECK(estl_server_wait(srv, a, 10) == ESTL_OK, "first done");
    Spec s{};
    s.kind = FFT;
    s.n = 2048;
    s.rows = 1;
    const Result first = collect(srv, s, a);
    hipStream_t st = nullptr;
    hipStreamCreate(&st);
    fake_stall_next_task_us(300000);
    CHECK(estl_server_copy_output(srv, a, dst.data(), 8 * 2048, st) == ESTL_OK, "copy_output");
    CHECK(estl_server_release(srv, a) == ESTL_OK, "release right after copy_output");
    CHECK(estl_server_submit_fft(srv, 2048, 1, ESTL_FORWARD, z2.data(), nullptr, &b) == ESTL_OK, "submit second");
    CHECK(estl_server_wait(srv, b, 10) == ESTL_OK, "second done");
    hipStreamSynchronize(st);
    CHECK(std::memcmp(dst.data(), first.out.data(), first.out.size()) == 0,
          "the copy queued before the release read another request's output");
    estl_server_release(srv, b);
    estl_server_destroy(srv);
    hipStreamDestroy(st);
    printf("%s copyout\n", g_failures ? "FAIL" : "PASS");
}

static void case_shutdown(unsigned seed) {
    std::mt19937_64 g(seed);
    auto filters = make_filters(g);
    auto reqs = make_stream(g, filters, 40);
    fake_set_stall_us(500);
    estl_server_t srv = nullptr;
    CHECK(estl_server_create(&srv, 8LL << 20, 3) == ESTL_OK, "create");
    std::vector<long long> fids;
    for (auto &f : filters) {
        long long id = 0;
        estl_server_register_filter(srv, f.n, f.h.data(), f.k, f.fp16 ? ESTL_CONV_FP16 : 0, nullptr, &id);
        fids.push_back(id);
    }
    int submitted = 0;
    for (auto &r : reqs) {
        long long id = 0;
        submitted += submit(srv, r, fids, &id) == ESTL_OK;
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    estl_server_destroy(srv);  // with requests queued and running
    fake_set_stall_us(0);
    printf("%s shutdown with %d requests submitted\n", g_failures ? "FAIL" : "PASS", submitted);
}

int main(int argc, char **argv) {
    const std::string which = argc > 1 ? argv[1] : "all";
    const unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 4;
    if (which == "all" || which == "mix") case_mix(seed);
    if (which == "all" || which == "denial") case_denial();
    if (which == "all" || which == "launchfail") case_launch_failure(seed + 1, false);
    if (which == "all" || which == "asyncfail") case_launch_failure(seed + 2, true);
    if (which == "all" || which == "stale") case_stale();
    if (which == "all" || which == "copyout") case_copy_output();
    if (which == "all" || which == "shutdown") case_shutdown(seed + 3);
    printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
    return g_failures ? 1 : 0;
}
