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
//   * request groups: a convolution (FP32 or FP16 storage) at the head of the queue
//     takes along the queued requests of the same kind, length and flags found in the
//     next kGroupWindow entries (up to kGroupMax requests, kGroupMaxValues values;
//     single FP32 rows, which run at half length, only with single rows), and the
//     group runs as one batch: one launch per pass for all of them (the pass kernels
//     find each row's request in a ConvGroup descriptor) instead of one chain of
//     launches per request.  Each member keeps its own output, id and release;
//   * resident state: plans, twiddle tables and convolution shapes are cached per size
//     and stay on the device; a filter can be registered once (its spectrum then lives
//     in the arena, counted against the budget) and used by any number of requests;
//   * the same pass split and kernels run for a request whatever runs beside it (a
//     group only changes which columns share a tile, not what a column computes), and
//     no value crosses requests, so each result is bit-identical alone and in the mix;
//   * latency (submit -> completion observed) and throughput are recorded per request;
//     completed ids are delivered through a completion queue.
// One dispatcher thread admits requests in arrival order; one launcher thread per stream
// (or fewer, ESTL_SERVER_LAUNCHERS) issues their HIP calls (outside the lock, streams in
// parallel); one completer thread watches the requests' events.
//
// Failure handling: a launch that fails after earlier launches of its request were
// queued leaves the request's memory in place until an event recorded after those
// launches completes (they may still write into it); only then does the request fail
// and its blocks return.  An output copy (estl_server_copy_output) records an event on
// the caller's stream; a release while such copies are queued returns the blocks once
// they have run.  The host logic is exercised on the CPU, under ThreadSanitizer and
// AddressSanitizer, over a stand-in runtime with injected launch failures
// (tests/host/).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
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
    // bytes held for the long term (registered filter spectra): a request larger than
    // size() - held() could never be admitted
    void hold(size_t bytes) { held_ += round(bytes); }
    void unhold(size_t bytes) { held_ -= round(bytes); }
    size_t held() const { return held_; }

  private:
    void *base_ = nullptr;
    static size_t round(size_t bytes) { return bytes ? (bytes + 255) & ~size_t(255) : 256; }
    size_t size_ = 0, used_ = 0, peak_ = 0, held_ = 0;
    std::map<size_t, size_t> free_;
};

enum ReqState { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };
enum ReqKind { KIND_FFT = 0, KIND_CONV = 1, KIND_FILTER = 2, KIND_CONV_REG = 3 };

struct Request {
    long long id = 0;
    int kind = KIND_FFT;
    long n = 0, rows = 0, k = 0;
    int dir = ESTL_FORWARD;
    unsigned flags = 0;
    const void *in = nullptr;
    const float *h = nullptr;
    const float *xscale = nullptr;  // FP16 storage (ESTL_CONV_FP16): per-row input scales
    long long filter = 0;  // KIND_CONV_REG: the registration it uses
    // convolutions: log2 of the overlap-save window, 0 for whole-row transforms
    // (conv_ols_log2 for the request's rows); registrations: nonzero when overlap-save
    int ols = 0;
    int users = 0;         // KIND_FILTER: requests queued or running on it
    // arena blocks: [0] output (the spectrum for KIND_FILTER), the rest workspace
    std::vector<std::pair<size_t, size_t>> blocks;
    void *out = nullptr;
    int state = QUEUED, error = ESTL_OK;
    std::string msg;
    Clock::time_point t_submit, t_done;
    hipEvent_t ev = nullptr;     // recorded after the request's last launch
    hipEvent_t ev_in = nullptr;  // recorded on the submitter's stream: inputs are ready
    // a launch failed after earlier launches of the request were queued: the request
    // fails with (error, msg) once `ev`, recorded after them, has completed
    bool launch_failed = false;
    // estl_server_copy_output copies still queued on callers' streams: the output block
    // goes back to the arena only once each of these has completed
    std::vector<hipEvent_t> copies;
    double alg_bytes = 0;
    // resolved at admission (the launcher threads never touch the caches)
    Plan *plan = nullptr;
    ConvShape *shape = nullptr;
    const float2 *spectrum = nullptr;
    int stream = 0;
    // request group: the leader launches and completes it; followers ride along
    std::vector<long long> followers;  // leader
    long long leader = 0;              // follower
    ConvShape *gshape = nullptr;       // leader: the group shape of (n, flags, pair bucket)
    long gpairs = 0;                   // leader: the group's pairs
    Plan *gspec = nullptr;             // leader: spectrum plan over the group's filters
    Plan *gfwd = nullptr, *ginv = nullptr;  // leader of single rows: half-length plans
    int gslot = -1;                    // leader: its pinned descriptor slot
};

// request groups
constexpr int kGroupWindow = 256;            // queued requests searched for companions
constexpr long kGroupMaxValues = 1L << 24;   // real values of a group
constexpr int kGroupSlots = 64;              // pinned descriptor slots (> in-flight groups)

class Server {
  public:
    int init(int device, long long budget, int nstreams) {
        device_ = device;
        nstreams_ = std::max(1, std::min(nstreams, 16));
        if (hipSetDevice(device) != hipSuccess) return set_error(ESTL_ERR_HIP, "hipSetDevice(%d)", device);
        cus_ = device_cus();
        int rc = arena_.init((size_t)budget);
        if (rc) return rc;
        for (int i = 0; i < nstreams_; ++i) {
            hipStream_t s;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return set_error(ESTL_ERR_HIP, "stream creation failed");
            streams_.push_back(s);
        }
        max_inflight_ = 2 * nstreams_;
        if (hipHostMalloc((void **)&host_desc_, sizeof(ConvGroup) * kGroupSlots) != hipSuccess)
            return set_error(ESTL_ERR_NOMEM, "pinned group descriptors");
        for (int i = kGroupSlots - 1; i >= 0; --i) free_slots_.push_back(i);
        // launcher threads: one per stream, or fewer when ESTL_SERVER_LAUNCHERS asks (launcher
        // i then issues the work of streams i, i + L, ...; the streams still run
        // concurrently on the device)
        nlaunchers_ = nstreams_;
        if (const char *e = getenv("ESTL_SERVER_LAUNCHERS")) {
            const int v = atoi(e);
            if (v >= 1) nlaunchers_ = std::min(v, nstreams_);
        }
        work_.resize(nstreams_);
        cv_work_ = std::vector<std::condition_variable>(nlaunchers_);
        for (int i = 0; i < nlaunchers_; ++i) launchers_.emplace_back([this, i] { launch_loop(i); });
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
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto &cv : cv_work_) cv.notify_all();
        }
        for (auto &t : launchers_)
            if (t.joinable()) t.join();
        if (completer_.joinable()) completer_.join();
        for (auto s : streams_) (void)hipStreamDestroy(s);
        for (auto &kv : reqs_) {
            if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
            if (kv.second.ev_in) (void)hipEventDestroy(kv.second.ev_in);
            for (hipEvent_t e : kv.second.copies) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
        }
        for (auto &d : deferred_)
            for (hipEvent_t e : d.copies) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
        for (auto e : events_) (void)hipEventDestroy(e);
        if (host_desc_) (void)hipHostFree(host_desc_);
    }

    int submit(Request r, hipStream_t input_stream, long long *id) {
        // the server's streams wait for the work already queued on the submitter's
        // stream: an event recorded there before the request is queued (outside the
        // lock: a runtime call the other threads need not wait for)
        {
            std::lock_guard<std::mutex> lk(mu_);
            int rc = get_event(&r.ev_in);
            if (rc) return rc;
        }
        if (hipEventRecord(r.ev_in, input_stream) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu_);
            put_event(r.ev_in);
            return set_error(ESTL_ERR_HIP, "event record on the input stream failed");
        }
        std::lock_guard<std::mutex> lk(mu_);
        auto give_back = [&](int code) {
            put_event(r.ev_in);
            return code;
        };
        if (stop_) return give_back(set_error(ESTL_ERR_INVALID, "server is shutting down"));
        if ((long)queue_.size() >= max_queue_) {
            ++busy_;
            return give_back(set_error(ESTL_ERR_BUSY, "request queue full (%ld)", max_queue_));
        }
        Request *filter = nullptr;
        if (r.kind == KIND_CONV_REG) {
            auto it = reqs_.find(r.filter);
            if (it == reqs_.end() || it->second.kind != KIND_FILTER)
                return give_back(set_error(ESTL_ERR_INVALID, "%lld is not a registered filter", r.filter));
            if (it->second.state == FAILED) return give_back(set_error(ESTL_ERR_INVALID, "filter %lld failed", r.filter));
            filter = &it->second;
            r.n = filter->n;
            r.k = filter->k;  // its spectrum's filter length (picks the pipeline)
            // the registration's whole-row choice and centring: an overlap-save spectrum
            // leaves the centring rotation to the convolution's output stores
            r.flags |= filter->flags & (ESTL_CONV_FULL_LENGTH | ESTL_CONV_CENTER);
            if ((filter->flags & ESTL_CONV_FP16) != (r.flags & ESTL_CONV_FP16))
                return give_back(
                    set_error(ESTL_ERR_INVALID, "filter %lld and the request differ in FP16 storage", r.filter));
        }
        // the window a call of this many rows takes alone (so a served result is bitwise the
        // single call's); a registration holds both spectra where both lengths are possible
        if (r.kind == KIND_CONV || (r.kind == KIND_CONV_REG && filter->ols))
            r.ols = conv_ols_log2(r.n, r.rows, r.k, r.flags, cus_);
        else if (r.kind == KIND_FILTER)
            r.ols = conv_ols_log2(r.n, 1, r.k, r.flags, cus_);
        if (filter) ++filter->users;
        r.id = ++next_id_;
        r.t_submit = Clock::now();
        if (r.kind != KIND_FILTER && first_submit_ == Clock::time_point()) first_submit_ = r.t_submit;
        *id = r.id;
        queue_.push_back(r.id);
        reqs_.emplace(r.id, std::move(r));
        if (reqs_.at(*id).kind != KIND_FILTER) ++submitted_;
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
        const auto deadline = Clock::now() + to_duration(timeout_s);
        while (true) {
            auto it = reqs_.find(id);
            if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
            if (it->second.state == DONE) return ESTL_OK;
            if (it->second.state == FAILED) return set_error(it->second.error, "%s", it->second.msg.c_str());
            if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout)
                return set_error(ESTL_ERR_BUSY, "request %lld not done within %.3f s", id, timeout_s);
        }
    }

    // ids of finished (DONE or FAILED) requests, oldest first; filter registrations are
    // not reported here
    int next_done(long long *ids, int max, double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto deadline = Clock::now() + to_duration(timeout_s);
        while (done_q_.empty())
            if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout) return 0;
        int c = 0;
        while (c < max && !done_q_.empty()) {
            ids[c++] = done_q_.front();
            done_q_.pop_front();
        }
        return c;
    }

    int copy_output(long long id, void *dst, long long bytes, hipStream_t st) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
        const Request &r = it->second;
        if (r.state != DONE) return set_error(ESTL_ERR_INVALID, "request %lld is not done", id);
        if (r.blocks.empty()) return set_error(ESTL_ERR_INVALID, "request %lld has no output", id);
        if (bytes < 0 || (size_t)bytes > r.blocks[0].second)
            return set_error(ESTL_ERR_INVALID, "copy of %lld bytes exceeds the output (%zu)", bytes, r.blocks[0].second);
        // the copy runs later on the caller's stream: an event after it keeps the output
        // block out of the arena until it has run, whenever the request is released
        hipEvent_t done = nullptr;
        if (int rc = get_event(&done)) return rc;
        if (hipMemcpyAsync(dst, r.out, (size_t)bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
            put_event(done);
            return set_error(ESTL_ERR_HIP, "output copy failed");
        }
        if (hipEventRecord(done, st) != hipSuccess) {
            // the copy is queued but cannot be marked: wait for it here
            (void)hipStreamSynchronize(st);
            put_event(done);
            return ESTL_OK;
        }
        it->second.copies.push_back(done);
        return ESTL_OK;
    }

    int release(long long id) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return set_error(ESTL_ERR_INVALID, "unknown request %lld", id);
        Request &r = it->second;
        if (r.state == QUEUED || r.state == RUNNING) return set_error(ESTL_ERR_BUSY, "request %lld still in progress", id);
        if (r.users) return set_error(ESTL_ERR_BUSY, "filter %lld is used by %d requests", id, r.users);
        if (r.kind == KIND_FILTER && r.state == DONE) arena_.unhold(r.blocks[0].second);
        bool copying = false;
        for (hipEvent_t e : r.copies) copying = copying || hipEventQuery(e) == hipErrorNotReady;
        if (copying) {
            // output copies still queued read these blocks: the completer returns them
            deferred_.push_back(Deferred{std::move(r.blocks), std::move(r.copies)});
        } else {
            for (auto &b : r.blocks) arena_.release(b.first, b.second);
            for (hipEvent_t &e : r.copies) put_event(e);
        }
        put_event(r.ev);
        put_event(r.ev_in);
        reqs_.erase(it);
        cv_dispatch_.notify_one();  // memory came back
        return ESTL_OK;
    }

    int drain(double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto deadline = Clock::now() + to_duration(timeout_s);
        while (!queue_.empty() || !inflight_.empty() || launching_)
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
                 "\"streams\": %d, \"groups\": %lld, \"grouped_requests\": %lld}",
                 submitted_, completed_, rejected_, failed_, busy_, queue_.size(), inflight_.size(), 1e3 * pct(50),
                 1e3 * pct(99), 1e3 * mean, lat.empty() ? -1.0 : 1e3 * lat.back(), elapsed,
                 elapsed > 0 ? completed_ / elapsed : 0.0, elapsed > 0 ? done_bytes_ / elapsed / 1e9 : 0.0,
                 arena_.size(), arena_.used(), arena_.peak(), peak_inflight_, nstreams_, groups_, grouped_);
        return buf;
    }

    void reset_stats() {
        std::lock_guard<std::mutex> lk(mu_);
        latencies_.clear();
        submitted_ = completed_ = rejected_ = failed_ = busy_ = groups_ = grouped_ = 0;
        done_bytes_ = 0;
        first_submit_ = last_done_ = Clock::time_point();
        peak_inflight_ = inflight_.size();
    }

  private:
    static Clock::duration to_duration(double s) {
        return std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(std::max(0.0, s)));
    }

    // Blocks of released requests whose output copies were still queued: returned to the
    // arena once every copy has run (completer thread).  True if any came back.
    bool return_deferred() {
        std::lock_guard<std::mutex> lk(mu_);
        bool any = false;
        for (auto it = deferred_.begin(); it != deferred_.end();) {
            bool pending = false;
            for (hipEvent_t e : it->copies) pending = pending || hipEventQuery(e) == hipErrorNotReady;
            if (pending) {
                ++it;
                continue;
            }
            for (auto &b : it->blocks) arena_.release(b.first, b.second);
            for (hipEvent_t &e : it->copies) put_event(e);
            it = deferred_.erase(it);
            any = true;
        }
        return any;
    }

    // event pool (under mu_)
    int get_event(hipEvent_t *e) {
        if (!events_.empty()) {
            *e = events_.back();
            events_.pop_back();
            return ESTL_OK;
        }
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
            return set_error(ESTL_ERR_HIP, "event creation failed");
        return ESTL_OK;
    }
    void put_event(hipEvent_t &e) {
        if (e) events_.push_back(e);
        e = nullptr;
    }

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
        const int prec = (flags & ESTL_CONV_FP16) ? ESTL_PRECISION_FP16 : ESTL_PRECISION_FP32;
        if (conv_shape_init(c.get(), n, rows, flags & ~(unsigned)ESTL_CONV_FP16, prec)) return nullptr;
        ConvShape *raw = c.get();
        conv_shapes_[key] = std::move(c);
        return raw;
    }
    static size_t spec_scratch(const ConvShape &c) {
        return c.spec_plan.kind == Plan::PASSES ? c.spec_plan.scratch_bytes : 0;
    }

    Plan *half_plan(long m, long batch, int dir) {
        auto key = std::make_tuple(m, batch, dir);
        auto it = half_plans_.find(key);
        if (it != half_plans_.end()) return it->second.get();
        std::unique_ptr<Plan> p(new Plan());
        if (build_plan(p.get(), m, batch, dir, ESTL_PRECISION_FP32, 0, nullptr)) return nullptr;
        if (p->kind != Plan::PASSES) return nullptr;
        Plan *raw = p.get();
        half_plans_[key] = std::move(p);
        return raw;
    }

    Plan *group_spec_plan(long n, long filters) {
        auto key = std::make_pair(n, filters);
        auto it = group_spec_plans_.find(key);
        if (it != group_spec_plans_.end()) return it->second.get();
        std::unique_ptr<Plan> p(new Plan());
        if (group_spectrum_plan(p.get(), n, filters)) return nullptr;
        Plan *raw = p.get();
        group_spec_plans_[key] = std::move(p);
        return raw;
    }

    // a single FP32 row of n >= 2^14, which runs through half-length transforms
    static bool half_row(const Request &r) {
        const int prec = (r.flags & ESTL_CONV_FP16) ? ESTL_PRECISION_FP16 : ESTL_PRECISION_FP32;
        return conv_half_path(r.n, r.rows, prec);
    }

    // a convolution whose filter is ready (inline, or registered and done); single rows
    // on the half-length path group with one another when their passes are global
    // (n/2 > 2^13) and their input can be read as complex values, as it is alone
    bool groupable(const Request &r) const {
        if (r.kind != KIND_CONV && r.kind != KIND_CONV_REG) return false;
        if (r.flags & ESTL_CONV_NORMALIZE_KERNEL) return false;
        if (r.ols) return false;  // one overlap-save launch alone
        if (half_row(r) && (r.n <= (2L << kMaxRowsLog2) || (reinterpret_cast<uintptr_t>(r.in) & 7))) return false;
        if (r.n < 2 || r.rows < 1) return false;
        if (r.kind == KIND_CONV_REG) {
            auto f = reqs_.find(r.filter);
            if (f == reqs_.end() || f->second.state != DONE) return false;
        }
        return true;
    }

    // Admission of the queue head `r` as the leader of a group (under the lock): its
    // companions from the next kGroupWindow queued requests, one output block each, and
    // the group's workspace on the leader: two scratch buffers, the spectra (inline
    // filters) and the device descriptor.  False (nothing changed) when there are no
    // companions or the group does not fit now; r is then admitted alone.
    bool admit_group(Request &r) {
        if (free_slots_.empty() || !groupable(r)) return false;
        std::vector<Request *> m{&r};
        long pairs = (r.rows + 1) / 2, values = r.n * r.rows;
        const size_t window = std::min(queue_.size(), (size_t)kGroupWindow);
        for (size_t i = 1; i < window && (int)m.size() < kGroupMax; ++i) {
            Request &o = reqs_.at(queue_[i]);
            if (o.kind != r.kind || o.n != r.n || o.flags != r.flags || half_row(o) != half_row(r) || !groupable(o))
                continue;
            if (values + o.n * o.rows > kGroupMaxValues) continue;
            m.push_back(&o);
            pairs += (o.rows + 1) / 2;
            values += o.n * o.rows;
        }
        if (m.size() < 2) return false;
        const long G = (long)m.size();
        const bool inline_filters = r.kind == KIND_CONV;
        // shapes for groups are cached per (n, flags, power of two of the pair count):
        // built for 2^b >= pairs pairs (tiles fit a batch within 2x of the group's) and
        // launched with the group's own pair count, so the cache stays bounded
        int bucket = 0;
        while ((1L << bucket) < pairs) ++bucket;
        ConvShape *gs = conv_shape(r.n, 2 * (1L << bucket), r.flags);
        if (!gs) return false;
        // multi-pass sizes: the spectra through a spectrum plan over the G filters
        Plan *gsp = (inline_filters && !gs->fused) ? group_spec_plan(r.n, G) : nullptr;
        if (inline_filters && !gs->fused && !gsp) return false;
        // single rows: half-length plans for the bucket's batch (launched with G rows)
        Plan *gfwd = nullptr, *ginv = nullptr;
        if (half_row(r)) {
            gfwd = half_plan(r.n / 2, 1L << bucket, ESTL_FORWARD);
            ginv = half_plan(r.n / 2, 1L << bucket, ESTL_BACKWARD);
            if (!gfwd || !ginv) return false;
        }
        // blocks: every member's output, then the leader's workspace (multi-pass: two
        // scratch buffers; inline filters: the spectra; the descriptor)
        const bool fp16 = (r.flags & ESTL_CONV_FP16) != 0;
        std::vector<size_t> sizes;
        for (Request *q : m)  // outputs as footprint() lays them out
            sizes.push_back(fp16 ? fp16_out_scales_offset(q->n, q->rows) + sizeof(float) * (size_t)q->rows
                                 : sizeof(float) * (size_t)q->n * (size_t)q->rows);
        if (!gs->fused) {  // the passes' ping-pong (FP32 or binary16 pairs), also the FP32 spectrum plan's
            const size_t scr = std::max((fp16 ? 4 : 8) * (size_t)r.n * (size_t)pairs,
                                        inline_filters ? sizeof(float2) * (size_t)r.n * (size_t)G : size_t(0));
            sizes.push_back(scr);
            sizes.push_back(scr);
        }
        if (inline_filters) sizes.push_back(sizeof(float2) * (size_t)r.n * (size_t)G);
        sizes.push_back(sizeof(ConvGroup));
        size_t total = 0;
        for (size_t b : sizes) total += (b + 255) & ~size_t(255);
        if (total > arena_.size() - arena_.held()) return false;
        std::vector<std::pair<size_t, size_t>> blocks;
        for (size_t b : sizes) {
            size_t off;
            if (!arena_.alloc(b, &off)) {
                for (auto &x : blocks) arena_.release(x.first, x.second);
                return false;
            }
            blocks.emplace_back(off, b);
        }
        hipEvent_t ev = nullptr;
        if (get_event(&ev)) {
            for (auto &x : blocks) arena_.release(x.first, x.second);
            return false;
        }
        // committed: out of the queue, outputs assigned, the leader carries the rest
        for (size_t i = 0; i < m.size(); ++i) {
            Request *q = m[i];
            queue_.erase(std::find(queue_.begin(), queue_.end(), q->id));
            q->blocks.assign(1, blocks[i]);
            q->out = arena_.base() + blocks[i].first;
            q->state = RUNNING;
            q->alg_bytes = (fp16 ? 4.0 : 8.0) * q->n * q->rows + (inline_filters ? 4.0 * q->k : 0.0);
            if (q->kind == KIND_CONV_REG) q->spectrum = static_cast<const float2 *>(reqs_.at(q->filter).out);
            if (i) {
                q->leader = r.id;
                r.followers.push_back(q->id);
            }
        }
        for (size_t i = m.size(); i < blocks.size(); ++i) r.blocks.push_back(blocks[i]);
        r.gshape = gs;
        r.gspec = gsp;
        r.gfwd = gfwd;
        r.ginv = ginv;
        r.gpairs = pairs;
        r.gslot = free_slots_.back();
        free_slots_.pop_back();
        r.ev = ev;
        r.stream = next_stream_++ % nstreams_;
        ++groups_;
        grouped_ += G;
        ++launching_;
        work_[r.stream].push_back(r.id);
        cv_work_[r.stream % nlaunchers_].notify_one();
        return true;
    }

    // The group's launches (launcher thread, no lock; `fol` resolved under the lock):
    // wait for every follower's inputs, upload the descriptor from the group's pinned
    // slot, the spectra of inline filters, then the convolution passes.
    int launch_group(Request &r, const std::vector<Request *> &fol, hipStream_t st) {
        for (Request *f : fol)
            if (hipStreamWaitEvent(st, f->ev_in, 0) != hipSuccess)
                return set_error(ESTL_ERR_HIP, "stream wait on a follower's input event failed");
        const bool inline_filters = r.kind == KIND_CONV;
        const bool fused = r.gshape->fused;
        const bool fp16 = (r.flags & ESTL_CONV_FP16) != 0;
        const long n = r.n;
        char *base = static_cast<char *>(arena_.base());
        // leader blocks: [0] output, then (multi-pass) two scratch buffers, (inline
        // filters) the spectra, and last the descriptor
        size_t next = 1;
        float2 *scr0 = nullptr, *scr1 = nullptr, *spectra = nullptr;
        if (!fused) {
            scr0 = reinterpret_cast<float2 *>(base + r.blocks[next++].first);
            scr1 = reinterpret_cast<float2 *>(base + r.blocks[next++].first);
        }
        if (inline_filters) spectra = reinterpret_cast<float2 *>(base + r.blocks[next++].first);
        if (next + 1 != r.blocks.size()) return set_error(ESTL_ERR_INVALID, "group workspace blocks");
        ConvGroup *d_desc = reinterpret_cast<ConvGroup *>(base + r.blocks.back().first);
        ConvGroup &g = host_desc_[r.gslot];
        memset(&g, 0, sizeof g);
        g.count = 1 + (int)fol.size();
        long pair = 0, max_rows = 0;
        for (int q = 0; q < g.count; ++q) {
            const Request &m = q ? *fol[q - 1] : r;
            if (m.n != n || m.rows < 1) return set_error(ESTL_ERR_INVALID, "group member %d does not match", q);
            g.pair_start[q] = pair;
            pair += (m.rows + 1) / 2;
            max_rows = std::max(max_rows, m.rows);
            g.real_rows[q] = m.rows;
            g.x[q] = m.in;
            g.y[q] = m.out;
            g.spec[q] = inline_filters ? spectra + (size_t)q * n : m.spectrum;
            g.h[q] = m.h;
            g.k[q] = m.k;
            g.shift[q] = (m.flags & ESTL_CONV_CENTER) ? ((m.k + 1) / 2) % n : 0;
            if (fp16) {
                g.xscale[q] = m.xscale;
                g.yscale[q] = reinterpret_cast<float *>(static_cast<char *>(m.out) + fp16_out_scales_offset(n, m.rows));
            }
            if (!g.x[q] || !g.y[q] || !g.spec[q] || (inline_filters && (!g.h[q] || g.k[q] < 1)) ||
                (fp16 && !g.xscale[q]))
                return set_error(ESTL_ERR_INVALID, "group member %d has a null buffer", q);
        }
        g.pair_start[g.count] = pair;
        if (pair != r.gpairs) return set_error(ESTL_ERR_INVALID, "group pairs %ld != admitted pairs %ld", pair, r.gpairs);
        if (hipMemcpyAsync(d_desc, &g, sizeof g, hipMemcpyHostToDevice, st) != hipSuccess)
            return set_error(ESTL_ERR_HIP, "group descriptor upload failed");
        // FP16 storage with inline filters: the members' L first, the spectra are stored
        // divided by it
        const bool divide_l1 = fp16 && inline_filters;
        if (divide_l1)
            if (int rc = conv_fp16_filter_group(d_desc, g.count, st)) return rc;
        if (inline_filters && fused) {
            if (int rc = conv_filter_group_rows(*r.gshape, d_desc, g.count, divide_l1, st)) return rc;
        } else if (inline_filters) {
            if (r.gspec->batch != g.count) return set_error(ESTL_ERR_INVALID, "group spectrum plan size");
            bool skip = true;  // the first spectrum pass is skipped when every filter allows it
            for (int q = 0; q < g.count; ++q) skip = skip && spectrum_skips_first_pass(*r.gspec, g.k[q], g.shift[q]);
            if (int rc = conv_filter_group(*r.gspec, d_desc, spectra, scr0, skip, divide_l1, st)) return rc;
        }
        if (r.gfwd)
            return conv_run_group_half(*r.gfwd, *r.ginv, d_desc, g.count, n, (r.flags & ESTL_CONV_CORRELATE) != 0, scr0,
                                       scr1, st);
        if (fp16) return conv_run_group_fp16(*r.gshape, d_desc, g.count, pair, max_rows, !inline_filters, scr0, scr1, st);
        return conv_run_group(*r.gshape, d_desc, pair, scr0, scr1, st);
    }

    // a group leader and its followers fail together; the slot comes back
    void fail_group_locked(Request &r, int code, const std::string &msg) {
        for (long long id : r.followers) {
            auto it = reqs_.find(id);
            if (it != reqs_.end()) fail_locked(it->second, code, msg);
        }
        if (r.gslot >= 0) free_slots_.push_back(r.gslot);
        r.gslot = -1;
        fail_locked(r, code, msg);
    }

    // sizes of the arena blocks a request needs: output first
    bool footprint(Request &r, std::vector<size_t> &sizes, std::string &why) {
        sizes.clear();
        auto fail = [&]() {
            why = estl_last_error();
            return false;
        };
        switch (r.kind) {
            case KIND_FFT: {
                Plan *p = fft_plan(r.n, r.rows, r.dir);
                if (!p) return fail();
                r.plan = p;
                sizes.push_back(sizeof(float2) * (size_t)r.n * (size_t)r.rows);
                if (p->kind == Plan::PASSES) sizes.push_back(p->scratch_bytes);
                r.alg_bytes = 16.0 * r.n * r.rows;
                return true;
            }
            case KIND_CONV:
            case KIND_CONV_REG: {
                ConvShape *c = conv_shape(r.n, r.rows, r.flags);
                if (!c) return fail();
                r.shape = c;
                if (r.flags & ESTL_CONV_FP16)  // binary16 rows, then (256-aligned) FP32 row scales
                    sizes.push_back(fp16_out_scales_offset(r.n, r.rows) + sizeof(float) * (size_t)r.rows);
                else
                    sizes.push_back(sizeof(float) * (size_t)r.n * (size_t)r.rows);
                // two scratch buffers: the pipeline's (multi-pass sizes) and, with an
                // inline filter, the spectrum transform's (it borrows the first); an
                // overlap-save request needs none, and a spectrum of its window length
                const size_t sb = request_scratch(r, *c);
                if (sb) {
                    sizes.push_back(sb);
                    sizes.push_back(sb);
                }
                if (r.kind == KIND_CONV) {
                    sizes.push_back(spectrum_bytes(r, *c));
                    sizes.push_back(256);
                }
                r.alg_bytes = ((r.flags & ESTL_CONV_FP16) ? 4.0 : 8.0) * r.n * r.rows + (r.kind == KIND_CONV ? 4.0 * r.k : 0.0);
                return true;
            }
            case KIND_FILTER: {
                ConvShape *c = conv_shape(r.n, 1, r.flags);
                if (!c) return fail();
                r.shape = c;
                sizes.push_back(spectrum_bytes(r, *c));
                sizes.push_back(256);
                if (spec_scratch(*c) && !r.ols) sizes.push_back(spec_scratch(*c));
                r.alg_bytes = 0;
                return true;
            }
        }
        why = "unknown request kind";
        return false;
    }

    // bytes of each scratch buffer of a convolution request (0: none)
    static size_t request_scratch(const Request &r, const ConvShape &c) {
        if (r.ols) return 0;
        return r.kind == KIND_CONV ? conv_scratch_bytes(c) : c.scratch_bytes;
    }
    // bytes of the filter spectrum an inline filter or a registration computes: n bins,
    // or an overlap-save window's (a registration: both windows' where both lengths are
    // possible); FP16 storage keeps one more value per spectrum (the L1 bound)
    static size_t spectrum_bytes(const Request &r, const ConvShape &c) {
        if (!r.ols) return c.spec_bytes;
        const long e = (r.flags & ESTL_CONV_FP16) ? 1 : 0;
        if (r.kind == KIND_FILTER && conv_ols_both(r.n, r.k, r.flags))
            return sizeof(float2) * (size_t)(conv_ols_wide_offset(e != 0) + (1L << kOlsWideLog2) + e);
        return sizeof(float2) * (size_t)((1L << r.ols) + e);
    }

    static size_t fp16_out_scales_offset(long n, long rows) {
        return (2 * (size_t)n * (size_t)rows + 255) & ~size_t(255);
    }

    // launches of one admitted request on stream st (called without the lock)
    int launch(Request &r, hipStream_t st) {
        auto ptr = [&](size_t i) -> float2 * { return reinterpret_cast<float2 *>(arena_.base() + r.blocks[i].first); };
        if (r.kind == KIND_FFT) {
            Plan *p = r.plan;
            float2 *ws0 = r.blocks.size() > 1 ? ptr(1) : nullptr;
            return execute_plan_ws(p, static_cast<const float2 *>(r.in), static_cast<float2 *>(r.out), ws0, nullptr, st);
        }
        if (r.kind == KIND_FILTER) {
            ConvShape *c = r.shape;
            ConvBuffers b;
            b.k = r.k;
            b.ols = r.ols;
            b.spec = ptr(0);
            b.dc = ptr(1);
            if (r.blocks.size() > 2) b.scratch[0] = ptr(2);
            if (r.ols && conv_ols_both(r.n, r.k, r.flags)) {
                // both window lengths: the 2^13 spectrum, then the 2^14 one after it
                b.ols = kOlsLog2;
                if (int rc = conv_filter(*c, b, r.h, r.k, st)) return rc;
                b.ols = kOlsWideLog2;
                b.spec += conv_ols_wide_offset((r.flags & ESTL_CONV_FP16) != 0);
            }
            return conv_filter(*c, b, r.h, r.k, st);
        }
        ConvShape *c = r.shape;
        ConvBuffers b;
        b.k = r.k;  // an inline filter's taps, or the registered filter's
        b.ols = r.ols;
        size_t next = 1;
        const size_t sb = request_scratch(r, *c);
        if (sb) {
            b.scratch[0] = ptr(next++);
            b.scratch[1] = ptr(next++);
        }
        if (r.kind == KIND_CONV) {
            b.spec = ptr(next++);
            b.dc = ptr(next++);
            int rc = conv_filter(*c, b, r.h, r.k, st);
            if (rc) return rc;
        } else {
            b.spec = const_cast<float2 *>(r.spectrum);
        }
        if (r.flags & ESTL_CONV_FP16) {
            char *o = static_cast<char *>(r.out);
            return conv_run_fp16(*c, b, r.in, r.xscale, o,
                                 reinterpret_cast<float *>(o + fp16_out_scales_offset(r.n, r.rows)), st);
        }
        return conv_run(*c, b, static_cast<const float *>(r.in), static_cast<float *>(r.out), st);
    }

    void finish_locked(Request &r) {
        if (r.kind == KIND_CONV_REG) {
            auto f = reqs_.find(r.filter);
            if (f != reqs_.end()) --f->second.users;
        }
        if (r.kind != KIND_FILTER) done_q_.push_back(r.id);
    }

    void fail_locked(Request &r, int code, const std::string &msg) {
        r.state = FAILED;
        r.error = code;
        r.msg = msg;
        for (auto &b : r.blocks) arena_.release(b.first, b.second);
        r.blocks.clear();
        r.out = nullptr;
        if (r.kind != KIND_FILTER) ++failed_;
        finish_locked(r);
        cv_done_.notify_all();
    }

    // Admission, in arrival order: footprint, arena blocks, a stream; the launch itself
    // is handed to that stream's launcher thread.
    void dispatch_loop() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_dispatch_.wait(lk, [&] {
                return stop_ || (!queue_.empty() && inflight_.size() + launching_ < (size_t)max_inflight_);
            });
            if (stop_) return;
            Request &r = reqs_.at(queue_.front());
            const Request *filter = nullptr;
            if (r.kind == KIND_CONV_REG) {
                auto f = reqs_.find(r.filter);  // held by r's use count
                filter = &f->second;
                if (filter->state == FAILED) {
                    queue_.pop_front();
                    fail_locked(r, ESTL_ERR_INVALID, "its filter registration failed");
                    continue;
                }
            }
            if (admit_group(r)) continue;
            std::vector<size_t> sizes;
            std::string why;
            if (!footprint(r, sizes, why)) {
                queue_.pop_front();
                fail_locked(r, ESTL_ERR_INVALID, why);
                continue;
            }
            size_t total = 0;
            for (size_t s : sizes) total += (s + 255) & ~size_t(255);
            if (total > arena_.size() - arena_.held()) {
                queue_.pop_front();
                if (r.kind != KIND_FILTER) ++rejected_;
                fail_locked(r, ESTL_ERR_NOMEM, "request needs " + std::to_string(total) + " bytes, the budget leaves " +
                                                   std::to_string(arena_.size() - arena_.held()));
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
            r.out = arena_.base() + r.blocks[0].first;
            r.state = RUNNING;
            if (r.kind == KIND_FILTER) arena_.hold(r.blocks[0].second);
            // a request on a filter still in flight goes to the filter's stream (stream
            // order makes the spectrum ready); others take turns
            if (filter) {
                r.spectrum = static_cast<const float2 *>(filter->out);
                if (r.ols == kOlsWideLog2 && conv_ols_both(filter->n, filter->k, filter->flags))
                    r.spectrum += conv_ols_wide_offset((filter->flags & ESTL_CONV_FP16) != 0);
                r.stream = filter->state == DONE ? next_stream_++ % nstreams_ : filter->stream;
            } else {
                r.stream = next_stream_++ % nstreams_;
            }
            const int rc = get_event(&r.ev);
            if (rc) {
                if (r.kind == KIND_FILTER) arena_.unhold(r.blocks[0].second);
                fail_locked(r, rc, estl_last_error());
                continue;
            }
            ++launching_;
            work_[r.stream].push_back(r.id);
            cv_work_[r.stream % nlaunchers_].notify_one();
        }
    }

    // Launcher i: the HIP calls of each admitted request of its streams (i, i + L, ...),
    // outside the lock, taking its streams in turn.
    void launch_loop(int li) {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        int turn = li;
        // this launcher's m streams li + j L, j < m, searched from `turn` on: one with
        // work, or -1
        const int m = (nstreams_ - li + nlaunchers_ - 1) / nlaunchers_;
        auto next_stream = [&]() -> int {
            const int t0 = (turn - li) / nlaunchers_;
            for (int j = 0; j < m; ++j) {
                const int s = li + ((t0 + j) % m) * nlaunchers_;
                if (!work_[s].empty()) return s;
            }
            return -1;
        };
        while (true) {
            int s = -1;
            // work first: admitted requests are launched even while stopping (the
            // completer waits for every one of them)
            cv_work_[li].wait(lk, [&] { return (s = next_stream()) >= 0 || stop_; });
            if (s < 0) return;  // stopping, nothing left to launch
            turn = s + nlaunchers_ < nstreams_ ? s + nlaunchers_ : li;
            const hipStream_t st = streams_[s];
            Request &r = reqs_.at(work_[s].front());
            work_[s].pop_front();
            std::vector<Request *> fol;  // a group's followers (running: not erased meanwhile)
            for (long long id : r.followers) fol.push_back(&reqs_.at(id));
            lk.unlock();
            int rc = ESTL_OK;
            if (hipStreamWaitEvent(st, r.ev_in, 0) != hipSuccess)
                rc = set_error(ESTL_ERR_HIP, "stream wait on the input event failed");
            if (!rc) rc = fol.empty() ? launch(r, st) : launch_group(r, fol, st);
            bool marked = false;  // r.ev follows everything of r queued on st
            if (!rc) {
                if (hipEventRecord(r.ev, st) == hipSuccess) marked = true;
                else rc = set_error(ESTL_ERR_HIP, "event record failed");
            }
            const std::string err = rc ? estl_last_error() : std::string();
            if (rc && !marked) {
                // launches of r queued before the failure may still run and write into
                // r's blocks (or read its descriptor slot): r keeps them until those
                // launches are done -- marked by an event, or else waited for here
                if (hipEventRecord(r.ev, st) == hipSuccess) marked = true;
                else (void)hipStreamSynchronize(st);
            }
            lk.lock();
            --launching_;
            if (rc && !marked) {
                if (r.kind == KIND_FILTER) arena_.unhold(r.blocks[0].second);
                if (fol.empty())
                    fail_locked(r, rc, err);
                else
                    fail_group_locked(r, rc, err);
                cv_dispatch_.notify_one();
                continue;
            }
            if (rc) {  // fails when r.ev completes (complete_loop)
                r.launch_failed = true;
                r.error = rc;
                r.msg = err;
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
                if (stop_ && inflight_.empty() && !launching_) return;
                watch.clear();
                for (long long id : inflight_) watch.emplace_back(id, reqs_.at(id).ev);
            }
            bool any = return_deferred();
            for (auto &w : watch) {
                const hipError_t q = hipEventQuery(w.second);
                if (q == hipErrorNotReady) continue;
                std::lock_guard<std::mutex> lk(mu_);
                Request &r = reqs_.at(w.first);
                inflight_.erase(std::find(inflight_.begin(), inflight_.end(), w.first));
                put_event(r.ev_in);
                if (q != hipSuccess || r.launch_failed) {
                    // a device error, or a launch failure whose earlier launches have now run
                    if (r.kind == KIND_FILTER) arena_.unhold(r.blocks[0].second);
                    const int code = r.launch_failed ? r.error : ESTL_ERR_HIP;
                    const std::string msg =
                        r.launch_failed ? r.msg : std::string("device error: ") + hipGetErrorString(q);
                    if (r.followers.empty())
                        fail_locked(r, code, msg);
                    else
                        fail_group_locked(r, code, msg);
                } else {
                    r.t_done = Clock::now();
                    r.state = DONE;
                    // workspace back to the arena now; the output stays until release
                    for (size_t i = 1; i < r.blocks.size(); ++i) arena_.release(r.blocks[i].first, r.blocks[i].second);
                    r.blocks.resize(1);
                    if (r.kind != KIND_FILTER) {
                        latencies_.push_back(std::chrono::duration<double>(r.t_done - r.t_submit).count());
                        ++completed_;
                        done_bytes_ += r.alg_bytes;
                        last_done_ = r.t_done;
                    }
                    finish_locked(r);
                    // a group completes with its leader
                    for (long long id : r.followers) {
                        Request &f = reqs_.at(id);
                        f.t_done = r.t_done;
                        f.state = DONE;
                        put_event(f.ev_in);
                        latencies_.push_back(std::chrono::duration<double>(f.t_done - f.t_submit).count());
                        ++completed_;
                        done_bytes_ += f.alg_bytes;
                        finish_locked(f);
                    }
                    r.followers.clear();
                    if (r.gslot >= 0) free_slots_.push_back(r.gslot);
                    r.gslot = -1;
                }
                any = true;
            }
            if (any) {
                cv_done_.notify_all();
                cv_dispatch_.notify_one();
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(5));
            }
        }
    }

    int device_ = 0, nstreams_ = 1, nlaunchers_ = 1, max_inflight_ = 2;
    int cus_ = 256;  // compute units (the overlap-save window choice)
    long max_queue_ = 1 << 16;
    Arena arena_;
    std::vector<hipStream_t> streams_;
    std::vector<hipEvent_t> events_;
    std::mutex mu_;
    std::condition_variable cv_dispatch_, cv_done_;
    std::deque<long long> queue_, done_q_;
    std::vector<long long> inflight_;
    std::map<long long, Request> reqs_;
    std::map<std::tuple<long, long, int>, std::unique_ptr<Plan>> fft_plans_;
    std::map<std::tuple<long, long, unsigned>, std::unique_ptr<ConvShape>> conv_shapes_;
    std::thread dispatcher_, completer_;
    std::vector<std::thread> launchers_;
    std::vector<std::deque<long long>> work_;  // admitted, to be launched on stream s
    std::vector<std::condition_variable> cv_work_;
    bool stop_ = false;
    size_t launching_ = 0;
    long long next_id_ = 0, submitted_ = 0, completed_ = 0, rejected_ = 0, failed_ = 0, busy_ = 0;
    size_t peak_inflight_ = 0;
    int next_stream_ = 0;
    double done_bytes_ = 0;
    std::vector<double> latencies_;
    Clock::time_point first_submit_, last_done_;
    // request groups
    std::map<std::pair<long, long>, std::unique_ptr<Plan>> group_spec_plans_;  // (n, filters)
    std::map<std::tuple<long, long, int>, std::unique_ptr<Plan>> half_plans_;  // (n / 2, batch, direction)
    ConvGroup *host_desc_ = nullptr;  // kGroupSlots pinned descriptors
    std::vector<int> free_slots_;
    struct Deferred {
        std::vector<std::pair<size_t, size_t>> blocks;
        std::vector<hipEvent_t> copies;
    };
    std::vector<Deferred> deferred_;  // released blocks waiting for their output copies
    long long groups_ = 0, grouped_ = 0;
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
    r.kind = KIND_FFT;
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
    r.kind = KIND_CONV;
    r.n = n;
    r.rows = rows;
    r.k = k;
    r.flags = conv_flags;
    r.in = d_x;
    r.h = d_h;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), request_id);
}

int estl_server_submit_conv_fp16(estl_server_t srv, long n, long rows, const void *d_x16, const float *d_x_scale,
                                 const float *d_h, long k, unsigned conv_flags, void *input_stream,
                                 long long *request_id) {
    if (!srv || !request_id || !d_x16 || !d_x_scale || !d_h) return set_error(ESTL_ERR_INVALID, "null argument");
    if (k < 1) return set_error(ESTL_ERR_INVALID, "filter length %ld < 1", k);
    Request r;
    r.kind = KIND_CONV;
    r.n = n;
    r.rows = rows;
    r.k = k;
    r.flags = conv_flags | ESTL_CONV_FP16;
    r.in = d_x16;
    r.xscale = d_x_scale;
    r.h = d_h;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), request_id);
}

int estl_server_submit_conv_filter_fp16(estl_server_t srv, long long filter_id, long rows, const void *d_x16,
                                        const float *d_x_scale, unsigned conv_flags, void *input_stream,
                                        long long *request_id) {
    if (!srv || !request_id || !d_x16 || !d_x_scale) return set_error(ESTL_ERR_INVALID, "null argument");
    if (conv_flags & ~(unsigned)ESTL_CONV_CORRELATE)
        return set_error(ESTL_ERR_INVALID, "only ESTL_CONV_CORRELATE is chosen per request (0x%x)", conv_flags);
    Request r;
    r.kind = KIND_CONV_REG;
    r.filter = filter_id;
    r.rows = rows;
    r.flags = conv_flags | ESTL_CONV_FP16;
    r.in = d_x16;
    r.xscale = d_x_scale;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), request_id);
}

int estl_server_register_filter(estl_server_t srv, long n, const float *d_h, long k, unsigned conv_flags,
                                void *input_stream, long long *filter_id) {
    if (!srv || !filter_id || !d_h) return set_error(ESTL_ERR_INVALID, "null argument");
    if (k < 1) return set_error(ESTL_ERR_INVALID, "filter length %ld < 1", k);
    if (conv_flags & ESTL_CONV_CORRELATE)
        return set_error(ESTL_ERR_INVALID, "correlation is chosen per request, not at registration");
    Request r;
    r.kind = KIND_FILTER;
    r.n = n;
    r.rows = 1;
    r.k = k;
    r.flags = conv_flags;
    r.h = d_h;
    r.in = d_h;
    return reinterpret_cast<Server *>(srv)->submit(std::move(r), static_cast<hipStream_t>(input_stream), filter_id);
}

int estl_server_submit_conv_filter(estl_server_t srv, long long filter_id, long rows, const float *d_x,
                                   unsigned conv_flags, void *input_stream, long long *request_id) {
    if (!srv || !request_id || !d_x) return set_error(ESTL_ERR_INVALID, "null argument");
    if (conv_flags & ~(unsigned)ESTL_CONV_CORRELATE)
        return set_error(ESTL_ERR_INVALID, "only ESTL_CONV_CORRELATE is chosen per request (0x%x)", conv_flags);
    Request r;
    r.kind = KIND_CONV_REG;
    r.filter = filter_id;
    r.rows = rows;
    r.flags = conv_flags;
    r.in = d_x;
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

int estl_server_next_done(estl_server_t srv, long long *request_ids, int max_ids, double timeout_s) {
    if (!srv || !request_ids || max_ids < 1) return set_error(ESTL_ERR_INVALID, "bad argument");
    return reinterpret_cast<Server *>(srv)->next_done(request_ids, max_ids, timeout_s);
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
