// fft_device.h -- device building blocks of the gfx950 FFT kernels.
//
// * complex helpers on float2 (interleaved complex64, the layout of the FFTW3 path,
//   FFT_FFTW3.h:24, and of ComplexArray(..., planar=false), Complex.h:156-192);
// * compile-time twiddle constants W_R^k = exp(DIR * 2 pi i k / R), evaluated in double
//   by constexpr series and rounded once to float (the reference printed them with six
//   significant digits, KernelGenerator.h:60-66; here they are correctly rounded);
// * register-resident DFTs of size 2..64 (natural order in and out) built from radix-4
//   and radix-2 decimation-in-time steps with the constant twiddles folded in.  They
//   play the role of OpenCLFFTAlgorithm::generateLocalFFTKernel (OpenCLFFTAlgorithm.h:14-43)
//   and of the +-1/+-i shortcuts of KomplexConstMultiplication (KernelGenerator.h:100-142),
//   as C++ templates instead of generated OpenCL strings.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace estl {

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * w
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}
// a * conj(w)
__device__ __forceinline__ float2 cmulc(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -a.x * w.y));
}
// a * w for DIR = -1 (w as tabulated, forward) or a * conj(w) for DIR = +1 (backward):
// tables hold the forward twiddles exp(-2 pi i m / M); the backward ones are their
// conjugates, which is exact.
template <int DIR>
__device__ __forceinline__ float2 ctw(float2 a, float2 w) {
    if constexpr (DIR < 0) return cmul(a, w);
    else return cmulc(a, w);
}

// Streaming (nontemporal) global accesses for data touched exactly once.
typedef float v2f_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2 *p) {
    v2f_t v = __builtin_nontemporal_load(reinterpret_cast<const v2f_t *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float2 *p, float2 x) {
    v2f_t v = {x.x, x.y};
    __builtin_nontemporal_store(v, reinterpret_cast<v2f_t *>(p));
}
__device__ __forceinline__ float ld_stream_f(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream_f(float *p, float x) { __builtin_nontemporal_store(x, p); }

// The reference's tile-planarized layout (PlannarizedComplexArray<5>, Complex.h:34-48,
// 97-130): in a float array, the real part of value e sits at ((e >> 5) << 6) | (e & 31)
// and its imaginary part 32 floats later -- 32 reals, then their 32 imaginaries.
__device__ __forceinline__ long planar_pos(long e) { return ((e >> 5) << 6) | (e & 31); }
__device__ __forceinline__ float2 ld_planar(const float2 *base, long e) {
    const float *f = reinterpret_cast<const float *>(base) + planar_pos(e);
    return make_float2(__builtin_nontemporal_load(f), __builtin_nontemporal_load(f + 32));
}
__device__ __forceinline__ void st_planar(float2 *base, long e, float2 v) {
    float *f = reinterpret_cast<float *>(base) + planar_pos(e);
    __builtin_nontemporal_store(v.x, f);
    __builtin_nontemporal_store(v.y, f + 32);
}

// FP16 storage of a complex value (re, im as IEEE binary16), computed on in FP32.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_elem(const float2 *p) { return ld_stream(p); }
__device__ __forceinline__ void st_elem(float2 *p, float2 x) { st_stream(p, x); }
__device__ __forceinline__ float2 ld_elem(const f16x2 *p) {
    const f16x2 h = __builtin_nontemporal_load(p);
    return make_float2((float)h.x, (float)h.y);
}
__device__ __forceinline__ void st_elem(f16x2 *p, float2 x) {
    const f16x2 h = {(_Float16)x.x, (_Float16)x.y};  // round to nearest even
    __builtin_nontemporal_store(h, p);
}

// Returns p unchanged, but hides from the optimiser that it is loop-invariant.  Used on
// twiddle-table pointers inside loops over rows/items: otherwise every twiddle load of a
// pass is hoisted out of the loop and ~30 float2 stay live for the whole kernel, which
// costs occupancy; re-issued per use they hit in L1/L2.
template <class T>
__device__ __forceinline__ const T *opaque(const T *p) {
    asm volatile("" : "+s"(p));
    return p;
}
// Same for a per-lane value (the thread index): in loops whose body dispatches on a
// size class, addresses derived from the thread index alone would otherwise be hoisted
// for every class at once and all stay live.
// Value e of a convolution filter of length n built from k real taps h: the taps folded
// modulo n (a filter longer than the signal wraps, the circular definition) and rotated
// by `shift` -- the sum of h[t] over t = e + shift (mod n), accumulated in FP64.
__device__ __forceinline__ float fold_tap(const float *__restrict__ h, long k, long n, long shift, long e) {
    double acc = 0.0;
    for (long t = (e + shift) % n; t < k; t += n) acc += (double)h[t];
    return (float)acc;
}

__device__ __forceinline__ int opaque_lane(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// ---------------------------------------------------------------- compile-time trig
namespace ct {
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double sin_series(double x) {  // |x| <= pi/4
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 16; ++i) {
        term *= -x2 / double((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double cos_series(double x) {  // |x| <= pi/4
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 16; ++i) {
        term *= -x2 / double((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
// cos and sin of 2 pi k / n for n a power of two: exact quarter-turn reduction, then a
// series on [-pi/4, pi/4].
constexpr long long pmod(long long k, long long n) { return ((k % n) + n) % n; }
constexpr double cos2pi(long long k, long long n) {
    long long m = pmod(k, n);
    if (n < 4) return (m == 0) ? 1.0 : -1.0;
    long long q = (4 * m + n / 2) / n;        // nearest quarter turn
    long long r = 4 * m - q * n;              // remainder in units of (2 pi / 4n)
    double phi = 2.0 * kPi * double(r) / double(4 * n);
    double c = cos_series(phi), s = sin_series(phi);
    switch (q & 3) {
        case 0: return c;
        case 1: return -s;
        case 2: return -c;
        default: return s;
    }
}
constexpr double sin2pi(long long k, long long n) {
    long long m = pmod(k, n);
    if (n < 4) return 0.0;
    long long q = (4 * m + n / 2) / n;
    long long r = 4 * m - q * n;
    double phi = 2.0 * kPi * double(r) / double(4 * n);
    double c = cos_series(phi), s = sin_series(phi);
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}
}  // namespace ct

// x * W_R^K with W_R = exp(DIR 2 pi i / R); trivial rotations become swaps and negations.
template <int K, int R, int DIR>
__device__ __forceinline__ float2 twc(float2 x) {
    constexpr int k = int(ct::pmod(K, R));
    if constexpr (k == 0) {
        return x;
    } else if constexpr ((4 * k) % R == 0) {
        constexpr int q = (4 * k) / R;           // quarter turns, 1..3
        constexpr int qd = (DIR < 0) ? (4 - q) % 4 : q;  // positive-direction turns
        if constexpr (qd == 1) return make_float2(-x.y, x.x);   // * i
        else if constexpr (qd == 2) return make_float2(-x.x, -x.y);
        else return make_float2(x.y, -x.x);                     // * -i
    } else if constexpr ((8 * k) % R == 0) {
        constexpr float h = 0.70710678118654752440f;
        constexpr float c = float(ct::cos2pi(k, R));
        constexpr float s = float(DIR) * float(ct::sin2pi(k, R));
        // c, s = +-h: two adds and two multiplies
        constexpr float sc = (c > 0) ? 1.0f : -1.0f;
        constexpr float ss = (s > 0) ? 1.0f : -1.0f;
        return make_float2(h * (sc * x.x - ss * x.y), h * (ss * x.x + sc * x.y));
    } else {
        constexpr float c = float(ct::cos2pi(k, R));
        constexpr float s = float(DIR) * float(ct::sin2pi(k, R));
        return make_float2(fmaf(x.x, c, -x.y * s), fmaf(x.x, s, x.y * c));
    }
}

// compile-time for loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E>
struct StaticFor {
    template <class F>
    static __device__ __forceinline__ void run(F &&f) {
        if constexpr (B < E) {
            f(std::integral_constant<int, B>{});
            StaticFor<B + 1, E>::run(f);
        }
    }
};

// ---------------------------------------------------------------- register DFTs
// In-place DFT of R elements x[0..R) (natural order), sign DIR, unnormalised.
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
    static __device__ __forceinline__ void run(float2 *) {}
};

template <int DIR>
struct Dft<2, DIR> {
    static __device__ __forceinline__ void run(float2 *x) {
        float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    static __device__ __forceinline__ void run(float2 *x) {
        float2 a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
        float2 a2 = cadd(x[1], x[3]), a3 = twc<1, 4, DIR>(csub(x[1], x[3]));
        x[0] = cadd(a0, a2);
        x[2] = csub(a0, a2);
        x[1] = cadd(a1, a3);
        x[3] = csub(a1, a3);
    }
};

// R = 4 M (R >= 8): y_r[m] = x[4 m + r], Y_r = DFT_M(y_r),
// X[k + M q] = sum_r W_R^{r k} Y_r[k] W_4^{r q}.
template <int R, int DIR>
struct Dft {
    static_assert(R >= 8 && (R & (R - 1)) == 0, "radix must be a power of two");
    static constexpr int M = R / 4;
    static __device__ __forceinline__ void run(float2 *x) {
        float2 y[4][M];
        StaticFor<0, 4>::run([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            StaticFor<0, M>::run([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                y[r][m] = x[4 * m + r];
            });
            Dft<M, DIR>::run(y[r]);
        });
        StaticFor<0, M>::run([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            float2 a0 = y[0][k];
            float2 a1 = twc<k, R, DIR>(y[1][k]);
            float2 a2 = twc<2 * k, R, DIR>(y[2][k]);
            float2 a3 = twc<3 * k, R, DIR>(y[3][k]);
            float2 b0 = cadd(a0, a2), b1 = csub(a0, a2);
            float2 b2 = cadd(a1, a3), b3 = twc<1, 4, DIR>(csub(a1, a3));
            x[k] = cadd(b0, b2);
            x[k + 2 * M] = csub(b0, b2);
            x[k + M] = cadd(b1, b3);
            x[k + 3 * M] = csub(b1, b3);
        });
    }
};

// ---------------------------------------------------------------- synthetic data
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace estl
