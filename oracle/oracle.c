/*
 * oracle.c -- CPU restatement of ESTL's FFT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker (or as the timed CPU baseline).
 *
 * What each function restates (reference paths are relative to the ESTL tree):
 *
 *   oracle_plan_radices   FFT_OpenCL_Contiguous.h:52-61 (global radix list: radix 16
 *                         passes, the 2^(log2N mod 4) remainder last) and
 *                         SimpleMath.h:7-11 (ceilog)
 *   oracle_prefix         OpenCLFFTAlgorithm.h:9-12 (computeParameters: A_q, N_q)
 *   stockham_pass_*       FFT_OpenCL_Contiguous.h:167-180 (one global pass: read
 *                         in[(L/N_q)(B_q g + s) + z], pre-twiddle W_{N_q}^{g s}, local
 *                         DFT_{B_q}, write out[(L/N_q)(g + A_q h) + z])
 *   local_dft_*           OpenCLFFTAlgorithm.h:14-43 (the in-register radix-2 Stockham
 *                         DFT of size B_q that every pass runs)
 *   oracle_fft64/32       FFT_OpenCL_Contiguous.h:97-142 (execute: ping-pong passes),
 *                         with FFTW's sign convention FFT_FFTW3.h:26-28 (forward = -1,
 *                         backward = +1, both unnormalised) applied to every batch row
 *   oracle_pass_states64  the per-pass state of the same algorithm, exposed so a test
 *                         can pin each pass against a brute-force DFT
 *   oracle_dft_ld         O(N^2) long-double DFT (brute force, no reference counterpart)
 *   oracle_conv_circular  the direct circular convolution that Convolution.h:95-148
 *                         computes through FFTs (transform, multiply, inverse, scale)
 *   oracle_planar_pos     Complex.h:40-44 (ArrayPlannarizer::getPosition with
 *                         GlobalPlannarLevel 5, elementLevel 1), the layout of
 *                         KernelGenerator.h:197-208
 *   oracle_fill_uniform   the seeded uniform[-1,1) generator the synthetic inputs use
 *                         (splitmix64 over the flat float index; no reference
 *                         counterpart -- the reference seeds rand() with time(NULL))
 *
 * Twiddles are evaluated in long double from an exactly reduced integer angle, so the
 * FP64 oracle carries no twiddle error beyond double rounding.  The FP32 variant rounds
 * those same twiddles to float (correctly rounded tables), unlike the reference's
 * native_cos/native_sin of a 6-digit 2*pi (KernelGenerator.h:48,64), whose bit-level
 * output cannot be reproduced on a CPU.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_MAX_PASSES 64

/* ceilog(n, 2) of SimpleMath.h:7-11: smallest r with 2^r >= n. */
static int ceilog2(long n) {
    int r = 0;
    for (long v = 1; v < n; v *= 2) r++;
    return r;
}

/* Global radix list of FFT_OpenCL_Contiguous.h:55-60 with radix 2^log2radix
 * (the reference fixes log2radix = 4).  Returns the pass count. */
int oracle_plan_radices(long n, int log2radix, int *radices) {
    int l2 = ceilog2(n), count = 0;
    if (log2radix < 1) log2radix = 4;
    for (int i = 0; i < l2 / log2radix; ++i) radices[count++] = 1 << log2radix;
    if (l2 % log2radix) radices[count++] = 1 << (l2 % log2radix);
    return count;
}

/* computeParameters of OpenCLFFTAlgorithm.h:9-12. */
void oracle_prefix(int passes, const int *radices, long *A, long *Nq) {
    long b = 1;
    for (int q = 0; q < passes; ++q) { A[q] = b; b *= radices[q]; Nq[q] = b; }
}

/* exp(sign * 2 pi i * k / n) evaluated in long double after exact reduction of k. */
static void unit_ld(long k, long n, int sign, long double *c, long double *s) {
    long m = k % n;
    if (m < 0) m += n;
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    *c = cosl(ang);
    *s = (long double)sign * sinl(ang);
    /* exact values at multiples of a quarter turn */
    if ((4 * m) % n == 0) {
        long quarter = (4 * m) / n;
        static const int cq[4] = {1, 0, -1, 0}, sq[4] = {0, 1, 0, -1};
        *c = cq[quarter & 3];
        *s = (long double)sign * sq[quarter & 3];
    }
}

/* ---------------- FP64 and FP32 Stockham restatement ---------------- */

/* Per-plan twiddles, evaluated once in long double and rounded to the working type:
 * pre[q][g*B_q + s] = W_{N_q}^{g s} for pass q, loc[k] = W_{Bmax}^{k}. */
#define DEFINE_STOCKHAM(REAL, SFX)                                                         \
typedef struct {                                                                           \
    int P, radices[ORACLE_MAX_PASSES], Bmax;                                               \
    long A[ORACLE_MAX_PASSES], Nq[ORACLE_MAX_PASSES];                                      \
    REAL *pre_r[ORACLE_MAX_PASSES], *pre_i[ORACLE_MAX_PASSES];                             \
    REAL *loc_r, *loc_i;                                                                   \
} plan_##SFX;                                                                              \
                                                                                           \
static void plan_free_##SFX(plan_##SFX *p) {                                               \
    for (int q = 0; q < p->P; ++q) { free(p->pre_r[q]); free(p->pre_i[q]); }               \
    free(p->loc_r); free(p->loc_i);                                                        \
}                                                                                          \
                                                                                           \
static int plan_make_##SFX(plan_##SFX *p, long n, int log2radix, int sign, int inject) {   \
    memset(p, 0, sizeof(*p));                                                              \
    p->P = oracle_plan_radices(n, log2radix, p->radices);                                  \
    oracle_prefix(p->P, p->radices, p->A, p->Nq);                                          \
    p->Bmax = 1;                                                                           \
    for (int q = 0; q < p->P; ++q) {                                                       \
        int B = p->radices[q];                                                             \
        if (B > 4096) { p->P = q; plan_free_##SFX(p); return -2; }                         \
        if (B > p->Bmax) p->Bmax = B;                                                      \
        p->pre_r[q] = (REAL *)malloc(sizeof(REAL) * p->Nq[q]);                             \
        p->pre_i[q] = (REAL *)malloc(sizeof(REAL) * p->Nq[q]);                             \
        for (long g = 0; g < p->A[q]; ++g)                                                 \
            for (long s = 0; s < B; ++s) {                                                 \
                long double c, sn;                                                         \
                unit_ld(g * s, p->Nq[q], sign, &c, &sn);                                   \
                if (inject == 1) { c = 1; sn = 0; }                                        \
                p->pre_r[q][g * B + s] = (REAL)c; p->pre_i[q][g * B + s] = (REAL)sn;       \
            }                                                                              \
    }                                                                                      \
    p->loc_r = (REAL *)malloc(sizeof(REAL) * p->Bmax);                                     \
    p->loc_i = (REAL *)malloc(sizeof(REAL) * p->Bmax);                                     \
    for (long k = 0; k < p->Bmax; ++k) {                                                   \
        long double c, sn;                                                                 \
        unit_ld(k, p->Bmax, sign, &c, &sn);                                                \
        if (inject == 2) sn = -sn;                                                         \
        p->loc_r[k] = (REAL)c; p->loc_i[k] = (REAL)sn;                                     \
    }                                                                                      \
    return 0;                                                                              \
}                                                                                          \
                                                                                           \
/* In-register local DFT of size B (OpenCLFFTAlgorithm.h:14-43): log2(B) radix-2          \
 * Stockham stages; stage (A_L, 2, N_L) computes target[(B/N_L)(g + A_L h) + z] =          \
 * sum_s W_2^{h s} W_{N_L}^{g s} source[(B/N_L)(2 g + s) + z]. */                          \
static void local_dft_##SFX(const plan_##SFX *p, int B, REAL *re, REAL *im, REAL *tr,      \
                            REAL *ti) {                                                    \
    REAL *sr = re, *si = im, *dr = tr, *di = ti;                                           \
    long AL = 1;                                                                           \
    while (AL < B) {                                                                       \
        long NL = AL * 2, S = B / NL, step = p->Bmax / NL;                                 \
        for (long g = 0; g < AL; ++g) {                                                    \
            REAL wr = p->loc_r[g * step], wi = p->loc_i[g * step];                         \
            for (long z = 0; z < S; ++z) {                                                 \
                REAL ar = sr[S * (2 * g) + z], ai = si[S * (2 * g) + z];                   \
                REAL b0r = sr[S * (2 * g + 1) + z], b0i = si[S * (2 * g + 1) + z];         \
                REAL br = b0r * wr - b0i * wi, bi = b0r * wi + b0i * wr;                   \
                dr[S * g + z] = ar + br;            di[S * g + z] = ai + bi;               \
                dr[S * (g + AL) + z] = ar - br;     di[S * (g + AL) + z] = ai - bi;        \
            }                                                                              \
        }                                                                                  \
        REAL *t;                                                                           \
        t = sr; sr = dr; dr = t;                                                           \
        t = si; si = di; di = t;                                                           \
        AL = NL;                                                                           \
    }                                                                                      \
    if (sr != re) { memcpy(re, sr, sizeof(REAL) * B); memcpy(im, si, sizeof(REAL) * B); }  \
}                                                                                          \
                                                                                           \
/* One global pass (FFT_OpenCL_Contiguous.h:167-180) over an interleaved row of length L.  \
 * inject 3 plants a wrong output stride (used only by the oracle's own tests). */         \
static void stockham_pass_##SFX(const plan_##SFX *p, int q, long L, const REAL *in,        \
                                REAL *out, int inject) {                                   \
    long A = p->A[q];                                                                      \
    int B = p->radices[q];                                                                 \
    long S = L / (A * B);                                                                  \
    REAL xr[4096], xi[4096], tr[4096], ti[4096];                                           \
    for (long j = 0; j < L / B; ++j) {                                                     \
        long g = j / S, z = j % S;                                                         \
        for (int s = 0; s < B; ++s) {                                                      \
            long idx = S * (B * g + s) + z;                                                \
            REAL vr = in[2 * idx], vi = in[2 * idx + 1];                                   \
            if (A > 1) {                                                                   \
                REAL wr = p->pre_r[q][g * B + s], wi = p->pre_i[q][g * B + s];             \
                REAL pr = vr * wr - vi * wi, pi = vr * wi + vi * wr;                       \
                vr = pr; vi = pi;                                                          \
            }                                                                              \
            xr[s] = vr; xi[s] = vi;                                                        \
        }                                                                                  \
        local_dft_##SFX(p, B, xr, xi, tr, ti);                                             \
        for (int h = 0; h < B; ++h) {                                                      \
            long o = (inject == 3) ? (S * g + z + h * A) : (S * (g + A * h) + z);          \
            out[2 * o] = xr[h]; out[2 * o + 1] = xi[h];                                    \
        }                                                                                  \
    }                                                                                      \
}

DEFINE_STOCKHAM(double, 64)
DEFINE_STOCKHAM(float, 32)

/* FFT of `batch` interleaved rows of length n (power of two), sign -1 forward, +1
 * backward, unnormalised.  Passes ping-pong between out and a scratch buffer as
 * FFT_OpenCL_Contiguous.h:108-127 does between its two device buffers.
 * inject: 0 faithful; 1 drops the pre-twiddles; 2 flips the local twiddle sign;
 * 3 writes with a wrong output stride (planted mistakes for the oracle's own tests). */
#define DEFINE_FFT(REAL, SFX)                                                              \
int oracle_fft##SFX##_ex(long n, long batch, const REAL *in, REAL *out, int sign,          \
                         int log2radix, int inject) {                                      \
    if (n < 1 || (n & (n - 1))) return -1;                                                 \
    if (n == 1) { memcpy(out, in, sizeof(REAL) * 2 * batch); return 0; }                   \
    plan_##SFX p;                                                                          \
    int rc = plan_make_##SFX(&p, n, log2radix, sign, inject);                              \
    if (rc) return rc;                                                                     \
    REAL *buf = (REAL *)malloc(sizeof(REAL) * 2 * n * 2);                                  \
    if (!buf) { plan_free_##SFX(&p); return -3; }                                          \
    for (long b = 0; b < batch; ++b) {                                                     \
        const REAL *src = in + 2 * n * b;                                                  \
        REAL *ping = buf, *pong = buf + 2 * n;                                             \
        for (int q = 0; q < p.P; ++q) {                                                    \
            REAL *dst = (q == p.P - 1) ? out + 2 * n * b : ((q & 1) ? pong : ping);        \
            stockham_pass_##SFX(&p, q, n, src, dst, inject);                               \
            src = dst;                                                                     \
        }                                                                                  \
    }                                                                                      \
    free(buf);                                                                             \
    plan_free_##SFX(&p);                                                                   \
    return 0;                                                                              \
}                                                                                          \
int oracle_fft##SFX(long n, long batch, const REAL *in, REAL *out, int sign) {             \
    return oracle_fft##SFX##_ex(n, batch, in, out, sign, 4, 0);                            \
}

DEFINE_FFT(double, 64)
DEFINE_FFT(float, 32)

/* State after each global pass q (0-based) of one row: states[q*2n .. ] interleaved.
 * After pass q, d_q[(L/N_q) f + z] = DFT_{N_q}(x[(L/N_q) t + z])[f]. */
int oracle_pass_states64(long n, const double *in, double *states, int sign, int log2radix) {
    if (n < 2 || (n & (n - 1))) return -1;
    plan_64 p;
    if (plan_make_64(&p, n, log2radix, sign, 0)) return -2;
    const double *src = in;
    for (int q = 0; q < p.P; ++q) {
        double *dst = states + 2 * n * q;
        stockham_pass_64(&p, q, n, src, dst, 0);
        src = dst;
    }
    int P = p.P;
    plan_free_64(&p);
    return P;
}

/* Brute-force DFT in long double, angle reduced exactly as (n*k) mod N. */
int oracle_dft_ld(long n, long batch, const double *in, double *out, int sign) {
    if (n < 1) return -1;
    long double *cr = (long double *)malloc(sizeof(long double) * n);
    long double *ci = (long double *)malloc(sizeof(long double) * n);
    if (!cr || !ci) { free(cr); free(ci); return -3; }
    for (long k = 0; k < n; ++k) unit_ld(k, n, sign, &cr[k], &ci[k]);
    for (long b = 0; b < batch; ++b) {
        const double *x = in + 2 * n * b;
        double *y = out + 2 * n * b;
        for (long k = 0; k < n; ++k) {
            long double ar = 0, ai = 0;
            for (long t = 0; t < n; ++t) {
                long m = (long)(((unsigned long)t * (unsigned long)k) % (unsigned long)n);
                ar += x[2 * t] * cr[m] - x[2 * t + 1] * ci[m];
                ai += x[2 * t] * ci[m] + x[2 * t + 1] * cr[m];
            }
            y[2 * k] = (double)ar; y[2 * k + 1] = (double)ai;
        }
    }
    free(cr); free(ci);
    return 0;
}

/* Circular convolution y[i] = sum_{t<K} h[t] x[(i - t) mod n] in long double
 * (a filter longer than n folds modulo n: the mathematical circular definition). */
int oracle_conv_circular(long n, long batch, const double *x, long K, const double *h,
                         double *y) {
    for (long b = 0; b < batch; ++b)
        for (long i = 0; i < n; ++i) {
            long double acc = 0;
            for (long t = 0; t < K; ++t) {
                long idx = (i - (t % n)) % n;
                if (idx < 0) idx += n;
                acc += (long double)h[t] * x[b * n + idx];
            }
            y[b * n + i] = (double)acc;
        }
    return 0;
}

/* ArrayPlannarizer::getPosition (Complex.h:42-44) at plannarLevel 5, elementLevel 1. */
long oracle_planar_pos(long index, int element) {
    return ((index >> 5) << 6) | ((long)element << 5) | (index & 31);
}

/* splitmix64 finaliser. */
static uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Synthetic uniform[-1,1) floats: out[i] = ((u >> 40) * 2^-23) - 1 with
 * u = splitmix64(splitmix64(seed) + offset + i).  Every value is exact in FP32. */
void oracle_fill_uniform(float *out, long count, uint64_t seed, uint64_t offset) {
    uint64_t key = splitmix64(seed);
    for (long i = 0; i < count; ++i) {
        uint64_t u = splitmix64(key + offset + (uint64_t)i);
        out[i] = (float)((double)(u >> 40) * (1.0 / 8388608.0) - 1.0);
    }
}
