// dropin_check.cpp -- FFT_HIP<float> driven through an FFT<D> interface exactly as the
// reference's harness drives its backends (tests/Test.cpp:45-184): a factory
// (FFTFactorySpecific, FFT.h:46-52), newFFT(n, forward, batch), setData / execute /
// getData, and FFT2D (FFT.h:54-84) over a factory.
//
// Built twice by oracle/build.py:
//   * against the ESTL tree's own FFT.h / DataInterface.h / Complex.h (compiled from the
//     files where they lie; output oracle/_ref/dropin_check) -- the drop-in proof;
//   * against include/estl's API-identical headers (output tests/cpp/dropin_check_own).
// Either binary takes no arguments, needs one GPU, prints one line per check and exits
// 0 when every check passes.
//
// Checks:
//   1. the known-answer test of TestFourierFloat (tests/Test.cpp:45-110): batch rows of
//      cos(2 pi k q / M) must transform to M/2 at bins k and M-k, |diff|^2 <= 1e-3 M;
//   2. the 2-D differential test of TestFourier2DFloat (tests/Test.cpp:113-184): FFT2D
//      over FFT_HIP against FFT2D over a naive DFT backend (double accumulation, exactly
//      reduced angles) -- the role the FFTW3 factory plays there -- |diff|^2 <= 1e-3 Mx My,
//      on seeded random input (the reference's input degenerates to a ramp);
//   3. a backward transform of the forward result gives M times the input on every
//      batch row (the reference's batched inverse conjugates row 0 only);
//   4. the per-kernel profiling report in the format of tests/Test.cpp:86-100;
//   1b. the known-answer test again through FFT_HIP_Planar, whose host arrays are
//      PlannarizedDataInterface tiles read and written by the kernels directly.
// The 2-D test also runs FFT2D_HIP (the device-resident 2-D transform) on the same
// input against the naive factory.
#include <FFT.h>
#include <FFT2D_HIP.h>
#include <FFT_HIP.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef float D;
typedef Complex<D> C;

// O(n^2) DFT backend for the differential test (double accumulation; angle 2 pi (k t
// mod n) / n reduced exactly in integers).
class FFT_NaiveDFT : public FFT<D> {
  public:
    static std::string getName() { return "naive DFT (double)"; }
    static int getMaxBatchSize(int) { return 1 << 16; }
    FFT_NaiveDFT(int n, bool iForward = true, int iBatchCount = 1) : FFT<D>(n, iBatchCount), forward(iForward) {
        dataInterface = arrays = new SplitInterleavedDataInterface<D>(n * iBatchCount, false);
    }
    virtual void execute() {
        const int n = size;
        const double sgn = forward ? -1.0 : 1.0;
        const D *in = arrays->in->getData();
        D *out = arrays->out->getData();
        for (int b = 0; b < batchCount; ++b)
            for (int k = 0; k < n; ++k) {
                double re = 0, im = 0;
                for (int t = 0; t < n; ++t) {
                    const long m = ((long)k * t) % n;
                    const double a = sgn * 2.0 * M_PI * (double)m / (double)n;
                    const double xr = in[2 * (b * n + t)], xi = in[2 * (b * n + t) + 1];
                    re += xr * cos(a) - xi * sin(a);
                    im += xr * sin(a) + xi * cos(a);
                }
                out[2 * (b * n + k)] = (D)re;
                out[2 * (b * n + k) + 1] = (D)im;
            }
    }

  private:
    bool forward;
    SplitInterleavedDataInterface<D> *arrays;
};

static int failures = 0;

static void report(bool ok, const char *what, double worst, double limit) {
    printf("%s %s (worst %.3g, limit %.3g)\n", ok ? "PASS" : "FAIL", what, worst, limit);
    if (!ok) ++failures;
}

static double norm2(C a) { return (double)a.getNormSquared(); }

int main() {
    FFTFactory<D> *hip = new FFTFactorySpecific<FFT_HIP<D> >();
    FFTFactory<D> *naive = new FFTFactorySpecific<FFT_NaiveDFT>();
    printf("backend: %s (max batch at n=4096: %d)\n", hip->getName().c_str(), hip->getMaxBatchSize(4096));

    // 1. known-answer test
    for (int b : {0, 2, 5}) {
        const int B = 1 << b;
        for (int m = 1; m <= 16; ++m) {
            const int M = 1 << m;
            if (B > hip->getMaxBatchSize(M)) continue;
            FFT<D> *fft = hip->newFFT(M, true, B);
            double worst = 0;
            for (int k = 1; 2 * k < M && k <= 3; ++k) {
                for (int bIx = 0; bIx < B; ++bIx) {
                    const int effK = ((k + bIx) % (M / 2 - 1 > 0 ? M / 2 - 1 : 1)) + 1;
                    for (int q = 0; q < M; ++q) fft->setData(q + bIx * M, C((D)cos(2 * M_PI * ((1.0 * effK * q) / M))));
                }
                fft->execute();
                for (int bIx = 0; bIx < B; ++bIx) {
                    const int effK = ((k + bIx) % (M / 2 - 1 > 0 ? M / 2 - 1 : 1)) + 1;
                    for (int q = 0; q < M; ++q) {
                        const C want = ((q == effK) || (q == M - effK)) ? C((D)(M / 2.0)) : C((D)0.0);
                        worst = std::max(worst, norm2(fft->getData(q + bIx * M) - want) / M);
                    }
                }
            }
            delete fft;
            if (M >= 4) {
                char what[96];
                snprintf(what, sizeof what, "known answer: M=%d batch=%d", M, B);
                report(worst <= 1e-3, what, worst, 1e-3);
            }
        }
    }

    // 1b. the same known-answer test through the planarized-tile backend (the host arrays
    //     are PlannarizedDataInterface tiles; the kernels read and write them directly)
    {
        FFTFactory<D> *planar = new FFTFactorySpecific<FFT_HIP_Planar<D> >();
        for (int b : {0, 3}) {
            const int B = 1 << b;
            for (int m = 2; m <= 16; m += 2) {
                const int M = 1 << m;
                FFT<D> *fft = planar->newFFT(M, true, B);
                double worst = 0;
                for (int bIx = 0; bIx < B; ++bIx) {
                    const int effK = (1 + bIx) % (M / 2 - 1 > 0 ? M / 2 - 1 : 1) + 1;
                    for (int q = 0; q < M; ++q) fft->setData(q + bIx * M, C((D)cos(2 * M_PI * ((1.0 * effK * q) / M))));
                }
                fft->execute();
                for (int bIx = 0; bIx < B; ++bIx) {
                    const int effK = (1 + bIx) % (M / 2 - 1 > 0 ? M / 2 - 1 : 1) + 1;
                    for (int q = 0; q < M; ++q) {
                        const C want = ((q == effK) || (q == M - effK)) ? C((D)(M / 2.0)) : C((D)0.0);
                        worst = std::max(worst, norm2(fft->getData(q + bIx * M) - want) / M);
                    }
                }
                delete fft;
                char what[96];
                snprintf(what, sizeof what, "known answer, planarized tiles: M=%d batch=%d", M, B);
                report(worst <= 1e-3, what, worst, 1e-3);
            }
        }
        delete planar;
    }

    // 2. 2-D differential test against the naive factory
    srand(12345);
    for (int mx = 0; mx <= 6; ++mx)
        for (int my = 0; my <= 6; ++my) {
            if (mx + my > 10) continue;
            const int Mx = 1 << mx, My = 1 << my;
            FFT2D<D> a(Mx, My, hip), r(Mx, My, naive);
            for (int q = 0; q < Mx * My; ++q) {
                const C v((D)(2.0 * rand() / RAND_MAX - 1.0), (D)(2.0 * rand() / RAND_MAX - 1.0));
                a.data[q] = v;
                r.data[q] = v;
            }
            FFT2D_HIP<D> d(Mx, My);
            for (int q = 0; q < Mx * My; ++q) d.data[q] = a.data[q];
            a.execute();
            r.execute();
            d.execute();
            double worst = 0, worst_d = 0;
            for (int q = 0; q < Mx * My; ++q) {
                worst = std::max(worst, norm2(a.data[q] - r.data[q]) / (Mx * My));
                worst_d = std::max(worst_d, norm2(d.data[q] - r.data[q]) / (Mx * My));
            }
            char what[96];
            snprintf(what, sizeof what, "FFT2D differential: %d x %d", Mx, My);
            report(worst <= 1e-3, what, worst, 1e-3);
            snprintf(what, sizeof what, "FFT2D_HIP (device 2-D) differential: %d x %d", Mx, My);
            report(worst_d <= 1e-3, what, worst_d, 1e-3);
        }

    // 3. batched inverse on every row
    {
        const int M = 1024, B = 7;
        FFT<D> *f = hip->newFFT(M, true, B), *g = hip->newFFT(M, false, B);
        std::vector<C> x(M * B);
        for (int i = 0; i < M * B; ++i) x[i] = C((D)(2.0 * rand() / RAND_MAX - 1.0), (D)(2.0 * rand() / RAND_MAX - 1.0));
        for (int i = 0; i < M * B; ++i) f->setData(i, x[i]);
        f->execute();
        for (int i = 0; i < M * B; ++i) g->setData(i, f->getData(i));
        g->execute();
        double worst = 0;
        for (int i = 0; i < M * B; ++i) worst = std::max(worst, norm2(g->getData(i) - (D)M * x[i]) / M);
        delete f;
        delete g;
        report(worst <= 1e-3, "batched inverse round trip: M=1024 batch=7, every row", worst, 1e-3);
    }

    // 4. the per-kernel report of tests/Test.cpp:86-100 from the profiling surface
    for (int m : {12, 16, 20}) {
        FFT_HIP<D> f(1 << m, true, m == 12 ? 4096 : 16);
        const bool ok = f.profile(5) && f.getKernelCount() > 0;
        double totalT = 0, totalOps = 0;
        for (int q = 0; q < f.getKernelCount(); ++q) totalT += f.getKernelTime(q);
        printf("  N=2^%d batch=%d:", m, f.batchCount);
        for (int q = 0; q < f.getKernelCount(); ++q) {
            const double t = f.getKernelTime(q), ops = f.getTotalComputationFlops(q);
            totalOps += ops;
            printf(" [K%d (%s): %.1f Gfs (%d%%)]", q + 1, f.getKernelInfo(q).c_str(), 1e-6 * ops / t,
                   (int)(100 * (t / totalT) + 0.5));
        }
        printf(" (Total: %.1f Gfs)\n", 1e-6 * totalOps / totalT);
        char what[64];
        snprintf(what, sizeof what, "profiling surface: N=2^%d", m);
        report(ok && totalT > 0 && totalOps == f.getTotalComputationFlops(), what, totalT, 1e9);
    }

    printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
