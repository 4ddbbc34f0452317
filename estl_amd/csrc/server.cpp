// server.cpp -- a request loop in front of the FFT and convolution pipelines.
//
// The reference has no serving layer: every FFT object is built, executed and freed by
// the caller (tests/Test.cpp:53-103), device memory is only checked by printing
// "Insufficient global memory" (FFT_OpenCL_Contiguous.h:49), and
// Convolution2D_FFT::convolve is a blocking call.  Here requests are served as they
// arrive:
//   * admission by HBM footprint: every request's device memory (its output and its
//     workspace) comes from one arena of a fixed budget; a request is admitted when its
//     footprint fits, waits (in arrival order) while earlier requests hold the memory,
//     and is rejected only if it could never fit the budget;
//   * continuous batching: admitted requests are launched at once on a pool of streams
//     and run concurrently with the ones already on the device; their memory is
//     returned the moment they complete, admitting the next;
//   * resident state: plans, twiddle tables and convolution shapes are cached per size
//     and stay on the device;
//   * the same plan and kernels run for a request whatever runs beside it, and no
//     value crosses requests, so each result is bit-identical alone and in the mix;
//   * latency (submit -> completion observed) and throughput are recorded per request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/estl_hip.h"
#include "plan.h"

namespace estl {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (offsets aligned to 256 bytes).
class Arena {
  public:
    int init(size_t bytes) {
        size_ = (bytes + 255) & ~size_t(255);
        if (hipMalloc(&base_, size_) != hipSuccess) return set_error(ESTL_ERR_NOMEM, "arena of %zu bytes", size_);
        free_[0] = size_;
        return ESTL_OK;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return static_cast<char *>(base_); }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

  private:
    void *base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

enum ReqState { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Request {
    long long id = 0;
    int kind = 0;  // 0 FFT, 1 convolution
    long n = 0, rows = 0, k = 0;
    int dir = ESTL_FORWARD;
    unsigned flags = 0;
    const void *in = nullptr;
    const float *h = nullptr;
    // arena blocks: [0] output, the rest workspace
    std::vector<std::pair<size_t, size_t>> blocks;
    void *out = nullptr;
    int state = QUEUED, error = ESTL_OK;
    std::string msg;
    Clock::time_point t_submit, t_done;
    hipEvent_t ev = nullptr;     // recorded after the request's last launch
    hipEvent_t ev_in = nullptr;  // recorded on the submitter's stream: inputs are ready
    double alg_bytes = 0;
};

class Server {
  public:
    int init(int device, long long budget, int nstreams) {
        device_ = device;
        nstreams_ = std::max(1, std::min(nstreams, 16));
        if (hipSetDevice(device) != hipSuccess) return set_error(ESTL_ERR_HIP, "hipSetDevice(%d)", device);
        int rc = arena_.init((size_t)budget);
        if (rc) return rc;
        for (int i = 0; i < nstreams_; ++i) {
            hipStream_t s;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return set_error(ESTL_ERR_HIP, "stream creation failed");
            streams_.push_back(s);
        }
        max_inflight_ = 2 * nstreams_;
        dispatcher_ = std::thread([this] { dispatch_loop(); });
        completer_ = std::thread([this] { complete_loop(); });
        return ESTL_OK;
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_dispatch_.notify_all();
        if (dispatcher_.joinable()) dispatcher_.join();
        if (completer_.joinable()) completer_.join();
        for (auto s : streams_) (void)hipStreamDestroy(s);
        for (auto &kv : reqs_) {
            if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
            if (kv.second.ev_in) (void)hipEventDestroy(kv.second.ev_in);
        }
    }

    int submit(Request r, hipStream_t input_stream, long long *id) {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) return set_error(ESTL_ERR_INVALID, "server is shutting down");
        if ((long)queue_.size() >= max_queue_) {
            ++busy_;
            return set_error(ESTL_ERR_BUSY, "request queue full (%ld)", max_queue_);
        }
        // the server's streams wait for the work already queued on the submitter's stream
        if (hipEventCreateWithFlags(&r.ev_in, hipEventDisableTiming) != hipSuccess)
            return set_error(ESTL_ERR_HIP, "event creation failed");
        if (hipEventRecord(r.ev_in, input_stream) != hipSuccess) {
            (void)hipEventDestroy(r.ev_in);
            return set_error(ESTL_ERR_HIP, "event record on the input stream failed");
        }
        r.id = ++next_id_;
        r.t_submit = Clock::now();
        if (first_submit_ == Clock::time_point()) first_submit_ = r.t_submit;
        *id = r.id;
        queue_.push_back(r.id);
        reqs_.emplace(r.id, std::move(r));
        ++submitted_;
        cv_dispatch_.notify_one();
        return ESTL_OK;
    }

    int poll(long long id, int *state, void **out, double *latency) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
        const Request &r = it->second;
        if (state) *state = r.state;
        if (out) *out = (r.state == DONE) ? r.out : nullptr;
        if (latency) *latency = (r.state == DONE) ? std::chrono::duration<double>(r.t_done - r.t_submit).count() : -1;
        if (r.state == FAILED) return set_error(r.error, "%s", r.msg.c_str());
        return ESTL_OK;
    }

    int wait(long long id, double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
        while (true) {
            auto it = reqs_.find(id);
            if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
            if (it->second.state == DONE) return ESTL_OK;
            if (it->second.state == FAILED) return set_error(it->second.error, "%s", it->second.msg.c_str());
            if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout)
                return set_error(ESTL_ERR_BUSY, "request %lld not done within %.3f s", id, timeout_s);
        }
    }

    int copy_output(long long id, void *dst, long long bytes, hipStream_t st) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
        const Request &r = it->second;
        if (r.state != DONE) return set_error(ESTL_ERR_INVALID, "request %lld is not done", id);
        if (bytes < 0 || (size_t)bytes > r.blocks[0].second)
            return set_error(ESTL_ERR_INVALID, "copy of %lld bytes exceeds the output (%zu)", bytes, r.blocks[0].second);
        if (hipMemcpyAsync(dst, r.out, (size_t)bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
            return set_error(ESTL_ERR_HIP, "output copy failed");
        return ESTL_OK;
    }

    int release(long long id) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
        if (it->second.state == QUEUED || it->second.state == RUNNING)
            return set_error(ESTL_ERR_BUSY, "request %lld still in progress", id);
        for (auto &b : it->second.blocks) arena_.release(b.first, b.second);
        if (it->second.ev) (void)hipEventDestroy(it->second.ev);
        if (it->second.ev_in) (void)hipEventDestroy(it->second.ev_in);
        reqs_.erase(it);
        cv_dispatch_.notify_one();  // memory came back
        return ESTL_OK;
    }

    int drain(double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
        while (!queue_.empty() || !inflight_.empty())
            if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout)
                return set_error(ESTL_ERR_BUSY, "drain timed out with %zu queued, %zu running", queue_.size(),
                                 inflight_.size());
        return ESTL_OK;
    }

    std::string stats() {
        std::lock_guard<std::mutex> lk(mu_);
        std::vector<double> lat = latencies_;
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            if (lat.empty()) return -1.0;
            const double k = (lat.size() - 1) * p / 100.0;
            const size_t lo = (size_t)k, hi = std::min(lat.size() - 1, lo + 1);
            return lat[lo] + (lat[hi] - lat[lo]) * (k - lo);
        };
        double mean = 0;
        for (double v : lat) mean += v;
        if (!lat.empty()) mean /= lat.size();
        const double elapsed = (completed_ && last_done_ > first_submit_)
                                   ? std::chrono::duration<double>(last_done_ - first_submit_).count()
                                   : 0.0;
        char buf[1024];
        snprintf(buf, sizeof buf,
                 "{\"submitted\": %lld, \"completed\": %lld, \"rejected\": %lld, \"failed\": %lld, \"busy\": %lld, "
                 "\"queued\": %zu, \"running\": %zu, \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"mean_ms\": %.4f, "
                 "\"max_ms\": %.4f, \"elapsed_s\": %.6f, \"throughput_req_s\": %.3f, \"throughput_gbs\": %.3f, "
                 "\"budget_bytes\": %zu, \"arena_used_bytes\": %zu, \"arena_peak_bytes\": %zu, \"peak_inflight\": %zu, "
                 "\"streams\": %d}",
                 submitted_, completed_, rejected_, failed_, busy_, queue_.size(), inflight_.size(), 1e3 * pct(50),
                 1e3 * pct(99), 1e3 * mean, lat.empty() ? -1.0 : 1e3 * lat.back(), elapsed,
                 elapsed > 0 ? completed_ / elapsed : 0.0, elapsed > 0 ? done_bytes_ / elapsed / 1e9 : 0.0,
                 arena_.size(), arena_.used(), arena_.peak(), peak_inflight_, nstreams_);
        return buf;
    }

    void reset_stats() {
        std::lock_guard<std::mutex> lk(mu_);
        latencies_.clear();
        submitted_ = completed_ = rejected_ = failed_ = busy_ = 0;
        done_bytes_ = 0;
        first_submit_ = last_done_ = Clock::time_point();
        peak_inflight_ = inflight_.size();
    }

  private:
    // plan / shape caches (dispatcher thread only)
    Plan *fft_plan(long n, long batch, int dir) {
        auto key = std::make_tuple(n, batch, dir);
        auto it = fft_plans_.find(key);
        if (it != fft_plans_.end()) return it->second.get();
        std::unique_ptr<Plan> p(new Plan());
        if (build_plan(p.get(), n, batch, dir, ESTL_PRECISION_FP32, 0, nullptr)) return nullptr;
        Plan *raw = p.get();
        fft_plans_[key] = std::move(p);
        return raw;
    }
    ConvShape *conv_shape(long n, long rows, unsigned flags) {
        auto key = std::make_tuple(n, rows, flags);
        auto it = conv_shapes_.find(key);
        if (it != conv_shapes_.end()) return it->second.get();
        std::unique_ptr<ConvShape> c(new ConvShape());
        if (conv_shape_init(c.get(), n, rows, flags)) return nullptr;
        ConvShape *raw = c.get();
        conv_shapes_[key] = std::move(c);
        return raw;
    }

    // sizes of the arena blocks a request needs: output first
    bool footprint(Request &r, std::vector<size_t> &sizes, std::string &why) {
        sizes.clear();
        if (r.kind == 0) {
            Plan *p = fft_plan(r.n, r.rows, r.dir);
            if (!p) {
                why = estl_last_error();
                return false;
            }
            sizes.push_back(sizeof(float2) * (size_t)r.n * (size_t)r.rows);
            if (p->kind == Plan::PASSES) sizes.push_back(p->scratch_bytes);
            r.alg_bytes = 16.0 * r.n * r.rows;
        } else {
            ConvShape *c = conv_shape(r.n, r.rows, r.flags);
            if (!c) {
                why = estl_last_error();
                return false;
            }
            sizes.push_back(sizeof(float) * (size_t)r.n * (size_t)r.rows);
            sizes.push_back(c->spec_bytes);
            sizes.push_back(c->spec_bytes);
            sizes.push_back(256);
            size_t sb = c->scratch_bytes ? c->scratch_bytes : (c->spec_plan.kind == Plan::PASSES ? c->spec_bytes : 0);
            if (sb) {
                sizes.push_back(sb);
                sizes.push_back(sb);
            }
            r.alg_bytes = 8.0 * r.n * r.rows + 4.0 * r.k;
        }
        return true;
    }

    int launch(Request &r, hipStream_t st) {
        auto ptr = [&](size_t i) -> float2 * { return reinterpret_cast<float2 *>(arena_.base() + r.blocks[i].first); };
        r.out = arena_.base() + r.blocks[0].first;
        if (r.kind == 0) {
            Plan *p = fft_plan(r.n, r.rows, r.dir);
            float2 *ws0 = r.blocks.size() > 1 ? ptr(1) : nullptr;
            return execute_plan_ws(p, static_cast<const float2 *>(r.in), static_cast<float2 *>(r.out), ws0, nullptr, st);
        }
        ConvShape *c = conv_shape(r.n, r.rows, r.flags);
        ConvBuffers b;
        b.spec = ptr(1);
        b.hp = ptr(2);
        b.dc = ptr(3);
        if (r.blocks.size() > 5) {
            b.scratch[0] = ptr(4);
            b.scratch[1] = ptr(5);
        }
        int rc = conv_filter(*c, b, r.h, r.k, st);
        if (!rc) rc = conv_run(*c, b, static_cast<const float *>(r.in), static_cast<float *>(r.out), st);
        return rc;
    }

    void fail_locked(Request &r, int code, const std::string &msg) {
        r.state = FAILED;
        r.error = code;
        r.msg = msg;
        for (auto &b : r.blocks) arena_.release(b.first, b.second);
        r.blocks.clear();
        ++failed_;
        cv_done_.notify_all();
    }

    void dispatch_loop() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_dispatch_.wait(lk, [&] { return stop_ || (!queue_.empty() && inflight_.size() < (size_t)max_inflight_); });
            if (stop_) return;
            Request &r = reqs_.at(queue_.front());
            std::vector<size_t> sizes;
            std::string why;
            if (!footprint(r, sizes, why)) {
                queue_.pop_front();
                fail_locked(r, ESTL_ERR_INVALID, why);
                continue;
            }
            size_t total = 0;
            for (size_t s : sizes) total += (s + 255) & ~size_t(255);
            if (total > arena_.size()) {
                queue_.pop_front();
                ++rejected_;
                fail_locked(r, ESTL_ERR_NOMEM, "request needs " + std::to_string(total) + " bytes, budget is " +
                                                   std::to_string(arena_.size()));
                continue;
            }
            std::vector<std::pair<size_t, size_t>> blocks;
            bool ok = true;
            for (size_t s : sizes) {
                size_t off;
                if (!arena_.alloc(s, &off)) {
                    ok = false;
                    break;
                }
                blocks.emplace_back(off, s);
            }
            if (!ok) {
                for (auto &b : blocks) arena_.release(b.first, b.second);
                // wait for memory: a completion or a release notifies
                cv_dispatch_.wait(lk);
                continue;
            }
            queue_.pop_front();
            r.blocks = std::move(blocks);
            r.state = RUNNING;
            const int s = next_stream_++ % nstreams_;
            int rc = ESTL_OK;
            if (hipStreamWaitEvent(streams_[s], r.ev_in, 0) != hipSuccess)
                rc = set_error(ESTL_ERR_HIP, "stream wait on the input event failed");
            if (!rc) rc = launch(r, streams_[s]);
            if (!rc && hipEventCreateWithFlags(&r.ev, hipEventDisableTiming) != hipSuccess)
                rc = set_error(ESTL_ERR_HIP, "event creation failed");
            if (!rc && hipEventRecord(r.ev, streams_[s]) != hipSuccess) rc = set_error(ESTL_ERR_HIP, "event record failed");
            if (rc) {
                fail_locked(r, rc, estl_last_error());
                continue;
            }
            inflight_.push_back(r.id);
            peak_inflight_ = std::max(peak_inflight_, inflight_.size());
        }
    }

    void complete_loop() {
        (void)hipSetDevice(device_);
        std::vector<std::pair<long long, hipEvent_t>> watch;
        while (true) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (stop_ && inflight_.empty()) return;
                watch.clear();
                for (long long id : inflight_) watch.emplace_back(id, reqs_.at(id).ev);
            }
            bool any = false;
            for (auto &w : watch) {
                const hipError_t q = hipEventQuery(w.second);
                if (q == hipErrorNotReady) continue;
                std::lock_guard<std::mutex> lk(mu_);
                Request &r = reqs_.at(w.first);
                inflight_.erase(std::find(inflight_.begin(), inflight_.end(), w.first));
                if (q != hipSuccess) {
                    fail_locked(r, ESTL_ERR_HIP, std::string("device error: ") + hipGetErrorString(q));
                } else {
                    r.t_done = Clock::now();
                    r.state = DONE;
                    // workspace back to the arena now; the output stays until release
                    for (size_t i = 1; i < r.blocks.size(); ++i) arena_.release(r.blocks[i].first, r.blocks[i].second);
                    r.blocks.resize(1);
                    latencies_.push_back(std::chrono::duration<double>(r.t_done - r.t_submit).count());
                    ++completed_;
                    done_bytes_ += r.alg_bytes;
                    last_done_ = r.t_done;
                }
                any = true;
            }
            if (any) {
                cv_done_.notify_all();
                cv_dispatch_.notify_one();
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(10));
            }
        }
    }

    int device_ = 0, nstreams_ = 1, max_inflight_ = 2;
    long max_queue_ = 1 << 16;
    Arena arena_;
    std::vector<hipStream_t> streams_;
    std::mutex mu_;
    std::condition_variable cv_dispatch_, cv_done_;
    std::deque<long long> queue_;
    std::vector<long long> inflight_;
    std::map<long long, Request> reqs_;
    std::map<std::tuple<long, long, int>, std::unique_ptr<Plan>> fft_plans_;
    std::map<std::tuple<long, long, unsigned>, std::unique_ptr<ConvShape>> conv_shapes_;
    std::thread dispatcher_, completer_;
    bool stop_ = false;
    long long next_id_ = 0, submitted_ = 0, completed_ = 0, rejected_ = 0, failed_ = 0, busy_ = 0;
    size_t peak_inflight_ = 0;
    int next_stream_ = 0;
    double done_bytes_ = 0;
    std::vector<double> latencies_;
    Clock::time_point first_submit_, last_done_;
};

}  // namespace estl

using namespace estl;

extern "C" {

int estl_server_create(estl_server_t *srv, long long hbm_budget_bytes, int streams) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle pointer is null");
    *srv = nullptr;
    if (hbm_budget_bytes < 4096) return set_error(ESTL_ERR_INVALID, "budget of %lld bytes is too small", hbm_budget_bytes);
    Server *s = new Server();
    int rc = s->init(current_device(), hbm_budget_bytes, streams);
    if (rc) {
        delete s;
        return rc;
    }
    *srv = reinterpret_cast<estl_server_t>(s);
    return ESTL_OK;
}

int estl_server_submit_fft(estl_server_t srv, long n, long batch, int direction, const void *d_in,
                           void *input_stream, long long *request_id) {
    if (!srv || !request_id || !d_in) return set_error(ESTL_ERR_INVALID, "null argument");
    Request r;
    r.kind = 0;
    r.n = n;
    r.rows = batch;
    r.dir = direction;
    r.in = d_in;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), request_id);
}

int estl_server_submit_conv(estl_server_t srv, long n, long rows, const float *d_x, const float *d_h, long k,
                            unsigned conv_flags, void *input_stream, long long *request_id) {
    if (!srv || !request_id || !d_x || !d_h) return set_error(ESTL_ERR_INVALID, "null argument");
    if (k < 1) return set_error(ESTL_ERR_INVALID, "filter length %ld < 1", k);
    Request r;
    r.kind = 1;
    r.n = n;
    r.rows = rows;
    r.k = k;
    r.flags = conv_flags;
    r.in = d_x;
    r.h = d_h;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), request_id);
}

int estl_server_poll(estl_server_t srv, long long request_id, int *state, void **d_out, double *latency_s) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return reinterpret_cast<Server *>(srv)->poll(request_id, state, d_out, latency_s);
}

int estl_server_wait(estl_server_t srv, long long request_id, double timeout_s) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return reinterpret_cast<Server *>(srv)->wait(request_id, timeout_s);
}

int estl_server_copy_output(estl_server_t srv, long long request_id, void *d_dst, long long bytes, void *stream) {
    if (!srv || !d_dst) return set_error(ESTL_ERR_INVALID, "null argument");
    return reinterpret_cast<Server *>(srv)->copy_output(request_id, d_dst, bytes, static_cast<hipStream_t>(stream));
}

int estl_server_release(estl_server_t srv, long long request_id) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return reinterpret_cast<Server *>(srv)->release(request_id);
}

int estl_server_drain(estl_server_t srv, double timeout_s) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    return reinterpret_cast<Server *>(srv)->drain(timeout_s);
}

int estl_server_stats(estl_server_t srv, char *json, int len) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    const std::string s = reinterpret_cast<Server *>(srv)->stats();
    if (json && len > 0) {
        strncpy(json, s.c_str(), (size_t)len - 1);
        json[len - 1] = 0;
    }
    return (int)s.size();
}

int estl_server_reset_stats(estl_server_t srv) {
    if (!srv) return set_error(ESTL_ERR_INVALID, "handle is null");
    reinterpret_cast<Server *>(srv)->reset_stats();
    return ESTL_OK;
}

int estl_server_destroy(estl_server_t srv) {
    delete reinterpret_cast<Server *>(srv);
    return ESTL_OK;
}

}  // extern "C"
