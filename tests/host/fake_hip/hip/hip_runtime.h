// hip/hip_runtime.h -- a host-only stand-in for the HIP runtime, for testing the
// library's host code (the request server) on the CPU under sanitizers.
//
// Test infrastructure only: it is on the include path of the host test build
// (tests/host/), never of the library.  It models what the server relies on:
//   * device memory is host memory (hipMalloc = aligned allocation);
//   * a stream is a FIFO of tasks run in order by one worker thread; the null stream
//     runs a task at once in the calling thread;
//   * an event records the position of the last task queued on a stream; it completes
//     when the worker reaches that position; hipStreamWaitEvent makes a stream wait for
//     the record that was current at the call; hipEventQuery reports completion or a
//     task failure injected since the previous record on that stream;
//   * hipMemcpyAsync is a task.
// Tasks are queued by the fake launch functions (tests/host/fake_launch.cpp).
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>

struct float2 {
    float x, y;
};

enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorLaunchFailure = 719,
};

enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};

#define hipStreamDefault 0x0u
#define hipStreamNonBlocking 0x1u
#define hipEventDefault 0x0u
#define hipEventDisableTiming 0x2u

namespace fakehip {

struct Event {
    std::mutex mu;
    std::condition_variable cv;
    uint64_t recorded = 0;   // generation of the latest record
    uint64_t completed = 0;  // generation the stream has reached
    hipError_t error = hipSuccess;
};

struct Stream {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> tasks;
    bool stop = false, busy = false;
    bool failed = false;  // a task failed since the last event record (consumed by it)
    std::thread worker;
    Stream() {
        worker = std::thread([this] { run(); });
    }
    ~Stream() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        worker.join();
    }
    void push(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu);
            tasks.push_back(std::move(f));
        }
        cv.notify_all();
    }
    void drain() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return tasks.empty() && !busy; });
    }
    void run() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv.wait(lk, [&] { return stop || !tasks.empty(); });
            if (tasks.empty()) return;
            std::function<void()> f = std::move(tasks.front());
            tasks.pop_front();
            busy = true;
            lk.unlock();
            f();
            lk.lock();
            busy = false;
            cv.notify_all();
        }
    }
};

// worker stall before each task (stall_us) and before the next task to start on any
// stream (stall_once_us, consumed by that task); set by the tests through fake_launch.cpp
inline std::atomic<int> &stall_us() {
    static std::atomic<int> v{0};
    return v;
}
inline std::atomic<int> &stall_once_us() {
    static std::atomic<int> v{0};
    return v;
}

// run f on stream s (the null stream: now, in this thread)
inline void enqueue(Stream *s, std::function<void()> f) {
    if (!s) {
        f();
        return;
    }
    s->push([f = std::move(f)] {
        const int us = stall_us().load() + stall_once_us().exchange(0);
        if (us > 0) std::this_thread::sleep_for(std::chrono::microseconds(us));
        f();
    });
}

}  // namespace fakehip

typedef fakehip::Stream *hipStream_t;
typedef fakehip::Event *hipEvent_t;

inline const char *hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "hipSuccess";
        case hipErrorInvalidValue: return "hipErrorInvalidValue";
        case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
        case hipErrorNotReady: return "hipErrorNotReady";
        case hipErrorLaunchFailure: return "hipErrorLaunchFailure (injected)";
    }
    return "unknown";
}

inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }

inline hipError_t hipMalloc(void **p, size_t bytes) {
    *p = std::aligned_alloc(256, ((bytes ? bytes : 1) + 255) & ~size_t(255));
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) {
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}
inline hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void **p, size_t bytes, unsigned = 0) { return hipMalloc(p, bytes); }
inline hipError_t hipHostFree(void *p) { return hipFree(p); }

inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = new fakehip::Stream();
    return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamSynchronize(hipStream_t s) {
    if (s) s->drain();
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
    if (s) {
        s->drain();
        delete s;
    }
    return hipSuccess;
}

inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
    *e = new fakehip::Event();
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}

inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr) {
    if (!e) return hipErrorInvalidValue;
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        gen = ++e->recorded;
    }
    auto mark = [e, gen, s] {
        hipError_t err = hipSuccess;
        if (s) {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->failed) err = hipErrorLaunchFailure;
            s->failed = false;
        }
        std::lock_guard<std::mutex> lk(e->mu);
        if (gen > e->completed) {
            e->completed = gen;
            e->error = err;
        }
        e->cv.notify_all();
    };
    fakehip::enqueue(s, mark);
    return hipSuccess;
}

inline hipError_t hipEventQuery(hipEvent_t e) {
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->completed < e->recorded) return hipErrorNotReady;
    return e->error;
}

inline hipError_t hipEventSynchronize(hipEvent_t e) {
    std::unique_lock<std::mutex> lk(e->mu);
    e->cv.wait(lk, [&] { return e->completed >= e->recorded; });
    return e->error;
}

inline hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    if (!e) return hipErrorInvalidValue;
    uint64_t target;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        target = e->recorded;
    }
    fakehip::enqueue(s, [e, target] {
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv.wait(lk, [&] { return e->completed >= target; });
    });
    return hipSuccess;
}

inline hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind, hipStream_t s = nullptr) {
    fakehip::enqueue(s, [dst, src, bytes] { std::memcpy(dst, src, bytes); });
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k) {
    return hipMemcpyAsync(dst, src, bytes, k, nullptr);
}
inline hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s = nullptr) {
    fakehip::enqueue(s, [dst, v, bytes] { std::memset(dst, v, bytes); });
    return hipSuccess;
}
