// limits_check.cpp -- host-only checks of the C++ drop-in's size limits (no GPU needed).
//
// The drop-in keeps the reference's int element indices (Complex.h:16-192,
// DataInterface.h:6-45), so a host array holds at most 2^31 scalars.  What is checked:
//   1. AlignedArray / ComplexArray / PlannarizedComplexArray throw std::length_error
//      for negative sizes and for sizes whose scalar count does not fit an int, instead
//      of wrapping (the reference computes 2 * size in int, Complex.h:160);
//   2. small arrays index exactly as the reference's layouts do;
//   3. FFT_HIP<float> and FFT_HIP_Planar<float> refuse n * batchCount above what the
//      host arrays hold before anything is allocated, and getMaxBatchSize caps the
//      batch there even when device memory (a 288 GB budget) would allow more;
//   4. the batch the cap returns constructs host arrays whose last element is
//      addressable (checked arithmetically, without allocating gigabytes).
// Built by tests/test_dropin_limits.py with AddressSanitizer and UBSan on the host;
// prints one line per check and exits 0 when all pass.
#include <FFT.h>
#include <FFT_HIP.h>

#include <climits>
#include <cstdio>
#include <stdexcept>

static int failures = 0;

#define CHECK(cond, what)                                                  \
    do {                                                                   \
        const bool ok_ = (cond);                                           \
        printf("%s %s\n", ok_ ? "PASS" : "FAIL", what);                    \
        if (!ok_) ++failures;                                              \
    } while (0)

template <class F>
static bool throws_length(F f) {
    try {
        f();
    } catch (const std::length_error &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

int main() {
    // 1. array constructors
    CHECK(throws_length([] { AlignedArray<float> a(-1); }), "AlignedArray(-1) throws length_error");
    CHECK(throws_length([] { AlignedArray<float> a(kEstlMaxScalars + 1); }),
          "AlignedArray(2^31 + 1) throws length_error");
    CHECK(throws_length([] { ComplexArray<float> a(-3, false); }), "ComplexArray(-3) throws length_error");
    CHECK(throws_length([] { ComplexArray<float> a(INT_MAX, false); }),
          "ComplexArray(INT_MAX) throws length_error (2 * size would wrap)");
    CHECK(throws_length([] { ComplexArray<float> a((1 << 30) + 1, true); }),
          "ComplexArray(2^30 + 1) throws length_error");
    CHECK(throws_length([] { PlannarizedComplexArray<5, float> a(1 << 30); }),
          "PlannarizedComplexArray(2^30) throws length_error (2^31 scalars)");
    CHECK(throws_length([] { PlannarizedComplexArray<5, float> a(-1); }),
          "PlannarizedComplexArray(-1) throws length_error");
    CHECK(!throws_length([] { ComplexArray<float> a(0, false); }), "ComplexArray(0) constructs");

    // 2. small arrays: the reference's layouts
    {
        ComplexArray<float> a(5, false);
        for (int k = 0; k < 5; ++k) a.setElement(k, Complex<float>(float(k), float(-k)));
        bool ok = true;
        for (int k = 0; k < 5; ++k) ok &= a.getData()[2 * k] == float(k) && a.getData()[2 * k + 1] == float(-k);
        CHECK(ok, "interleaved ComplexArray(5) stores (re, im) pairs");
        ComplexArray<float> p(5, true);
        for (int k = 0; k < 5; ++k) p.setElement(k, Complex<float>(float(k), float(10 + k)));
        ok = true;
        for (int k = 0; k < 5; ++k) ok &= p.getReals()[k] == float(k) && p.getImaginaries()[k] == float(10 + k);
        CHECK(ok, "split ComplexArray(5) stores reals then imaginaries");
        PlannarizedComplexArray<5, float> t(70);
        CHECK(t.getElements() == 3 * 64, "PlannarizedComplexArray(70) holds 3 tiles of 64 scalars");
        t.setElement(69, Complex<float>(1.5f, -2.5f));
        CHECK(t.getData()[2 * 64 + 5] == 1.5f && t.getData()[2 * 64 + 32 + 5] == -2.5f,
              "value 69 lands in tile 2, slot 5 (real) and 37 (imaginary)");
    }

    // 3. FFT_HIP / FFT_HIP_Planar size checks (thrown before any allocation or device call)
    CHECK(throws_length([] { FFT_HIP<float> f(4096, true, 1 << 18); }),
          "FFT_HIP(4096, batch 2^18 = 2^30 values) throws length_error");
    CHECK(throws_length([] { FFT_HIP<float> f(1 << 30, true, 1); }), "FFT_HIP(2^30) throws length_error");
    CHECK(throws_length([] { FFT_HIP<float> f(4096, true, 0); }), "FFT_HIP(4096, batch 0) throws length_error");
    CHECK(throws_length([] { FFT_HIP<float> f(4096, true, INT_MAX); }),
          "FFT_HIP(4096, batch INT_MAX) throws length_error (n * batch wraps in int)");
    CHECK(throws_length([] { FFT_HIP_Planar<float> f(1 << 29, true, 2); }),
          "FFT_HIP_Planar(2^29, batch 2) throws length_error");

    const long long budget = 288LL * 1000 * 1000 * 1000;  // 288 GB of HBM
    const int b4096 = estl_dropin_max_batch(4096, kEstlDropinMaxValues, budget);
    printf("     max batch at n=4096, 288 GB: device %lld, drop-in %d\n",
           estl_max_batch(4096, ESTL_PRECISION_FP32, budget), b4096);
    CHECK(estl_max_batch(4096, ESTL_PRECISION_FP32, budget) > (1 << 20),
          "the device budget alone allows more than 2^20 rows of 4096");
    CHECK(b4096 == (INT_MAX / 2) / 4096, "FFT_HIP caps the batch at (INT_MAX / 2) / n");
    CHECK(estl_dropin_max_batch(1 << 26, kEstlDropinMaxValues, budget) == 15,
          "n = 2^26: the cap (15 rows) is below the budget (178 rows)");
    CHECK(estl_dropin_max_batch(1 << 30, kEstlDropinMaxValues, budget) == 1, "n = 2^30: at least 1 is reported");
    CHECK(estl_dropin_max_batch(64, kEstlDropinMaxValues, 1000) == 1, "a tiny budget reports 1");
    const int bp = estl_dropin_max_batch(4096, kEstlMaxPlanarValues, budget);
    CHECK(bp == (int)(kEstlMaxPlanarValues / 4096), "FFT_HIP_Planar caps the batch at (2^30 - 32) / n");

    // 4. the capped batch is addressable: last scalar index fits an int
    {
        const long long v = 4096LL * b4096;
        CHECK(2 * v - 1 <= INT_MAX, "interleaved: last scalar of n * max batch fits an int");
        const long long tiles = (4096LL * bp + 31) / 32;
        CHECK(tiles * 64 <= INT_MAX, "planarized: scalar count of n * max batch fits an int");
    }

    printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
