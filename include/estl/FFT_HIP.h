// FFT_HIP.h -- the gfx950 backend behind the reference's FFT<D> interface.
//
// Drop-in beside FFT_FFTW3 / FFT_OpenCL_Contiguous: same constructor (n, forward,
// batchCount), static getName() and getMaxBatchSize(n) (what FFTFactorySpecific calls,
// FFT.h:46-52), interleaved complex64 host arrays through SplitInterleavedDataInterface
// (planar = false), exactly the data interface of FFT_FFTW3 (FFT_FFTW3.h:24), and a
// synchronous execute() (copy in, compute, copy out, wait) like
// FFT_OpenCL_Contiguous::execute (FFT_OpenCL_Contiguous.h:97-142).  The work goes
// through the C ABI of estl_hip.h; nothing of HIP leaks into this header.
//
// Semantics are FFTW's (FFT_FFTW3.h:26-28): forward = exp(-2 pi i t k / n), backward =
// exp(+...), both unnormalised, applied to every batch row.  Setup failures (size not
// a power of two, no device, out of memory) throw std::runtime_error, and a size whose
// host arrays an int cannot index (n * batchCount of 2^30 complex values or more, or
// non-positive) throws std::length_error before anything is allocated; a failure inside
// execute() is printed and leaves the output untouched, the reference's behaviour
// (FFT_OpenCL_Contiguous.h:120-122).
#ifndef ESTL_HIP_FFT_HIP_H
#define ESTL_HIP_FFT_HIP_H

#include <climits>
#include <cstdio>
#include <stdexcept>
#include <string>

// Angle brackets on purpose: with the ESTL tree ahead of include/estl on the include
// path this compiles against ESTL's own FFT.h / DataInterface.h / Complex.h; without
// it, against the API-identical copies in include/estl.  Needs include/ on the path.
#include <FFT.h>
#include <estl_hip.h>

// largest batch of length-n rows the device's free memory holds (64-bit query; the
// reference's int query overflows on large devices, uOpenCL.h:323-337), capped so that
// n * batch host values stay within `max_values` (what the host arrays can index);
// budget_bytes = 0 asks the device, a positive value stands for its free memory
inline int estl_dropin_max_batch(int n, long long max_values, long long budget_bytes = 0) {
    if (n < 1) return 1;
    long long b = estl_max_batch(n, ESTL_PRECISION_FP32, budget_bytes);
    const long long cap = max_values / n;
    if (b > cap) b = cap;
    if (b > INT_MAX) b = INT_MAX;
    return b < 1 ? 1 : int(b);
}

// n * batchCount values as an int, or std::length_error if the host arrays cannot hold them
inline int estl_dropin_values(const char *who, int n, int batchCount, long long max_values) {
    const long long v = (long long)n * batchCount;
    if (n < 1 || batchCount < 1 || v > max_values)
        throw std::length_error(std::string(who) + ": " + std::to_string(n) + " x " + std::to_string(batchCount) +
                                " values do not fit host arrays indexed by int (at most " +
                                std::to_string(max_values) + ")");
    return int(v);
}

// Host-array limits that hold with this project's Complex.h and with the reference's
// (whose ComplexArray allocates 2 * size in int, Complex.h:160): an interleaved array
// holds at most INT_MAX / 2 = 2^30 - 1 values; a PlannarizedComplexArray rounds up to
// tiles of 32 values and counts its scalars in an int, so at most 2^30 - 32 values.
static const long long kEstlDropinMaxValues = INT_MAX / 2;
static const long long kEstlMaxPlanarValues = ((long long)INT_MAX / 64) * 32;

template <class D>
class FFT_HIP;

template <>
class FFT_HIP<float> : public FFT<float> {
  public:
    typedef SplitInterleavedDataInterface<float> DataInterfaceType;

    static std::string getName() { return "HIP FFT (gfx950)"; }

    // FFT_OpenCL_Contiguous.h:43-44, FFT.h:30
    static int getMaxBatchSize(int n) { return estl_dropin_max_batch(n, kEstlDropinMaxValues); }

    FFT_HIP(int iSize, bool iForward = true, int iBatchCount = 1, unsigned flags = 0)
        : FFT<float>(iSize, iBatchCount), forward(iForward), arrays(nullptr), plan(nullptr) {
        const int values = estl_dropin_values("FFT_HIP", iSize, iBatchCount, kEstlDropinMaxValues);
        this->dataInterface = arrays = new DataInterfaceType(values, false);
        const int rc = estl_plan_create(&plan, iSize, iBatchCount, iForward ? ESTL_FORWARD : ESTL_BACKWARD,
                                        ESTL_PRECISION_FP32, flags);
        if (rc != ESTL_OK) throw std::runtime_error(std::string("FFT_HIP: ") + estl_last_error());
    }

    virtual ~FFT_HIP() { estl_plan_destroy(plan); }

    virtual void execute() {
        const int rc = estl_execute_host(plan, arrays->in->getData(), arrays->out->getData());
        if (rc != ESTL_OK) printf("FFT_HIP::execute failed: %s\n", estl_last_error());
    }

    // profiling surface of the reference's OpenCL backends (kernelTimers,
    // getTotalComputationFlops(q), getKernelInfo(q): FFT_OpenCL_Contiguous.h:73-78,94-95)
    // as tests/Test.cpp:86-100 prints it: profile() times every launch on the plan's
    // resident buffers, then getKernelTime(q) holds its mean in milliseconds
    int getKernelCount() const { return estl_plan_launches(plan); }
    double getTotalComputationFlops() const {
        int l = 0;
        while ((1 << l) < size) ++l;
        return 5.0 * size * l * batchCount;
    }
    double getTotalComputationFlops(int q) const { return estl_plan_kernel_flops(plan, q); }
    std::string getKernelInfo(int q) const {
        char buf[256];
        return estl_plan_kernel_info(plan, q, buf, sizeof buf) < 0 ? std::string() : std::string(buf);
    }
    double getKernelTime(int q) const { return estl_plan_kernel_ms(plan, q); }
    bool profile(int reps = 10) { return estl_plan_profile(plan, nullptr, nullptr, reps, nullptr, 0) >= 0; }
    std::string describe() const {
        char buf[2048];
        estl_plan_describe(plan, buf, sizeof buf);
        return buf;
    }
    estl_plan_t getPlan() const { return plan; }

  private:
    bool forward;
    DataInterfaceType *arrays;
    estl_plan_t plan;
};

// The same backend behind PlannarizedDataInterface (DataInterface.h:26-35): the host
// arrays are PlannarizedComplexArray<GlobalPlannarLevel = 5> tiles (32 reals, then their
// 32 imaginaries), and the kernels read and write that layout directly
// (ESTL_FLAG_PLANAR): no repacking on the host or the device.
template <class D>
class FFT_HIP_Planar;

template <>
class FFT_HIP_Planar<float> : public FFT<float> {
  public:
    typedef PlannarizedDataInterface<float> DataInterfaceType;

    static std::string getName() { return "HIP FFT (gfx950, planarized tiles)"; }
    static int getMaxBatchSize(int n) { return estl_dropin_max_batch(n, kEstlMaxPlanarValues); }

    FFT_HIP_Planar(int iSize, bool iForward = true, int iBatchCount = 1, unsigned flags = 0)
        : FFT<float>(iSize, iBatchCount), arrays(nullptr), plan(nullptr) {
        const int values = estl_dropin_values("FFT_HIP_Planar", iSize, iBatchCount, kEstlMaxPlanarValues);
        this->dataInterface = arrays = new DataInterfaceType(values);
        const int rc = estl_plan_create(&plan, iSize, iBatchCount, iForward ? ESTL_FORWARD : ESTL_BACKWARD,
                                        ESTL_PRECISION_FP32, flags | ESTL_FLAG_PLANAR);
        if (rc != ESTL_OK) throw std::runtime_error(std::string("FFT_HIP_Planar: ") + estl_last_error());
    }

    virtual ~FFT_HIP_Planar() { estl_plan_destroy(plan); }

    virtual void execute() {
        const int rc = estl_execute_host(plan, arrays->in->getData(), arrays->out->getData());
        if (rc != ESTL_OK) printf("FFT_HIP_Planar::execute failed: %s\n", estl_last_error());
    }

    int getKernelCount() const { return estl_plan_launches(plan); }

  private:
    DataInterfaceType *arrays;
    estl_plan_t plan;
};

#endif  // ESTL_HIP_FFT_HIP_H
