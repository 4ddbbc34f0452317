// conv.hip -- the transform-multiply-inverse pipeline of Convolution.h:95-148, batched
// over real 1-D rows and fused on the device.
//
// The reference pads, transforms data and kernel with FFT2D (one FFT object and one
// host round trip per row and column, FFT.h:65-81), multiplies on the host
// (Convolution.h:125-137), transforms back and scales by 1/(K^[0] W H)
// (Convolution.h:141-143).  Here:
//   * the length-n filter is h folded modulo n (a filter longer than the signal
//                   wraps, the circular definition), optionally rotated so its centre
//                   (K + 1)/2 lands on index 0 (the reference's kernel shift,
//                   Convolution.h:113-120), built by the spectrum transform as it loads
//                   the taps (fold_tap): for n <= 2^13 k_filter_spectrum (fold, forward
//                   FFT and the kernel-sum normalisation in one workgroup), above that
//                   the first pass of the spectrum plan; the spectrum stays resident;
//   * k_conv_rows   for n <= 2^13: two real rows enter as one complex row (re = row 2p,
//                   im = row 2p + 1 -- the filter is real, so the two convolutions stay
//                   separate), forward FFT, multiply by the spectrum (its conjugate for
//                   correlation, Convolution.h:125-131), inverse FFT, scale, and the
//                   re / im parts go back to the two rows: one HBM read and one write
//                   per value, one launch for the whole batch;
//   * larger n runs the same pipeline as global passes with fused I/O (pass_tile.h);
//   * k_conv_rows on binary16 rows, k_filter_l1, k_spec_l1, k_conv_scales: the FP16-storage
//                   variant (and group forms of the last three for request groups).
#include "fft_block.h"
#include "internal.h"

namespace estl {

// sum |h_t| over k filter taps by the first 256 threads of a workgroup in a fixed
// summation order (thread t adds taps t, t + 256, ...; then a tree over the 256
// partial sums), so every caller gets the same value; every thread of the workgroup
// must call it (it holds barriers); the result is valid in every thread.
__device__ __forceinline__ float filter_l1_256(const float *__restrict__ h, long k, float *part) {
    if (threadIdx.x < 256) {
        float acc = 0.0f;
        for (long i = threadIdx.x; i < k; i += 256) acc += fabsf(h[i]);
        part[threadIdx.x] = acc;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    return part[0];
}

// spectrum /= Re spectrum[0] (the kernel sum), read from a copy `dc` taken beforehand
__global__ void k_spec_normalize(float2 *__restrict__ spec, long n, const float2 *__restrict__ dc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = 1.0f / dc->x;
    spec[i] = make_float2(spec[i].x * s, spec[i].y * s);
}

hipError_t launch_spec_normalize(float2 *spec, long n, const float2 *dc, hipStream_t st) {
    hipLaunchKernelGGL(k_spec_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spec, n, dc);
    return hipGetLastError();
}

// The member of a request group that owns pair `pair` (pass_tile.h has the same rule).
__device__ __forceinline__ int conv_group_member(const ConvGroup *g, long pair) {
    int q = 0;
    while (q + 1 < g->count && g->pair_start[q + 1] <= pair) ++q;
    return q;
}

// The spectrum of a filter of length n <= 2^13 in one workgroup: fold the taps on
// load, forward FFT (the rows kernel's block_fft and table, so the bins are those of the
// rows plan), optionally divide by the kernel sum Re K^[0] (Convolution.h:141), store.
// With a request group, workgroup q builds member q's filter into its spectrum, divided
// by the member's L (group->l1[q].x) for FP16 storage when divide_l1 is set.  With
// normalisation the kernel sum K^[0] is also written to dc (the FP16 L1 bound divides
// by it).
template <int LOG2N>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT)
k_filter_spectrum(const float *__restrict__ h, long k, long shift, float2 *__restrict__ spec,
                  const float2 *__restrict__ tw, int normalize, const ConvGroup *__restrict__ group, int divide_l1,
                  float2 *__restrict__ dc, float2 *__restrict__ l1_single) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    __shared__ float2 lds[G::uses_lds ? Lds<kLdsPad>::size(N) : 1];
    __shared__ float s_dc;
    __shared__ float s_part[TPT >= 256 ? 256 : 1];
    const int tl = opaque_lane(threadIdx.x);
    if (group) {
        const int q = blockIdx.x;
        h = group->h[q];
        k = group->k[q];
        shift = group->shift[q];
        spec = const_cast<float2 *>(group->spec[q]);
    }
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = make_float2(fold_tap(h, k, N, shift, tl + TPT * i), 0.0f);
    // kSpecTw: the power table below 4096 points, the block table from 4096
    block_fft<LOG2N, -1, kLdsPad, false, kSpecTw<LOG2N>>(v, lds, tl, tw, false);
    if (normalize) {
        if (tl == 0) {
            s_dc = v[0].x;  // bin 0
            if (dc) *dc = v[0];
        }
        __syncthreads();
        const float s = 1.0f / s_dc;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = make_float2(v[i].x * s, v[i].y * s);
    }
    if (group && divide_l1) {
        const float s = 1.0f / group->l1[blockIdx.x].x;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = make_float2(v[i].x * s, v[i].y * s);
    }
    if constexpr (TPT >= 256) {
        if (!group && l1_single) {
            // FP16 storage, one filter: L = sum |h| (divided by |kernel sum| when
            // normalised), the spectrum times 1 / L, and (L, sum |h|) after it -- the
            // values of k_filter_l1 then k_spec_l1, in one launch
            const float l1 = filter_l1_256(h, k, s_part);
            float L = l1;
            if (normalize) L /= fabsf(s_dc);
            if (!(L > 0.0f)) L = 1.0f;
            const float s = 1.0f / L;
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = make_float2(v[i].x * s, v[i].y * s);
            if (tl == 0) *l1_single = make_float2(L, l1);
        }
    }
#pragma unroll
    for (int i = 0; i < E; ++i) spec[tl + TPT * i] = v[i];
}

// l1_single (one filter, FP16 storage, N >= 2^12): the spectrum divided by its L and
// (L, sum |h|) stored there, in place of launch_conv_fp16_filter's two launches
hipError_t launch_filter_spectrum(int log2n, const float *h, long k, long shift, float2 *spec, const float2 *tw,
                                  bool normalize, hipStream_t st, const ConvGroup *group, int members,
                                  bool divide_l1, float2 *dc, float2 *l1_single) {
    if (group && (members < 1 || members > kGroupMax || normalize)) return hipErrorInvalidValue;
    if (l1_single && (group || log2n < 12)) return hipErrorInvalidValue;
    const unsigned blocks = group ? (unsigned)members : 1u;
    switch (log2n) {
#define FS(L)                                                                                                   \
    case L:                                                                                                     \
        hipLaunchKernelGGL(k_filter_spectrum<L>, dim3(blocks), dim3(BlockGeom<L>::TPT), 0, st, h, k, shift, spec, \
                           tw, normalize ? 1 : 0, group, divide_l1 ? 1 : 0, dc, l1_single);                      \
        break;
        FS(1) FS(2) FS(3) FS(4) FS(5) FS(6) FS(7) FS(8) FS(9) FS(10) FS(11) FS(12) FS(13) FS(14)
#undef FS
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// real rows stored as FP32 (streaming loads / stores) or binary16 (FP32 arithmetic)
__device__ __forceinline__ float ld_real(const float *p) { return ld_stream_f(p); }
__device__ __forceinline__ float ld_real(const _Float16 *p) { return (float)*p; }
__device__ __forceinline__ void st_real(float *p, float v) { st_stream_f(p, v); }
__device__ __forceinline__ void st_real(_Float16 *p, float v) { *p = (_Float16)v; }

// Persistent over pairs of real rows (TPW pairs per workgroup, four resident rounds of
// workgroups; one workgroup per group measured within 2 % either way at 1024-8192
// points, so the grid stays).  With
// a request group, `rows` counts the group's pairs twice and each slot takes its pair's
// member: input rows, spectrum and output rows.  R: the stored real type, float or
// _Float16 (FP16 storage: the spectrum has been divided by the filter's L1 norm, so
// every stored output lies in [-1, 1] (|y| <= sum |h| max |x|); the caller's per-row
// scales carry the magnitudes).
template <int LOG2N, int TPW, bool CORR, typename R>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT * TPW) __attribute__((amdgpu_waves_per_eu(4)))
k_conv_rows(const R *__restrict__ x, R *__restrict__ y, long rows, const float2 *__restrict__ spec,
            const float2 *__restrict__ tw, float scale, const ConvGroup *__restrict__ group) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    constexpr int SLICE = G::uses_lds ? Lds<kLdsPad>::size(N) : 0;
    constexpr int LDS = G::uses_lds ? TPW * SLICE : 1;
    __shared__ float2 lds[LDS];
    const long pairs = (rows + 1) / 2;
    const long groups = (pairs + TPW - 1) / TPW;
    for (long grp = blockIdx.x; grp < groups; grp += gridDim.x) {  // uniform across the workgroup
        const int t = opaque_lane(threadIdx.x);
        const int w = t / TPT, tl = t % TPT;
        const long pr = grp * TPW + w;
        const bool active = pr < pairs;
        long r1 = active ? 2 * pr : 0, nrows = rows;
        if (group) {
            const int q = active ? conv_group_member(group, pr) : 0;
            r1 = active ? 2 * (pr - group->pair_start[q]) : 0;
            nrows = group->real_rows[q];
            x = static_cast<const R *>(group->x[q]);
            y = static_cast<R *>(group->y[q]);
            spec = group->spec[q];
        }
        const bool two = active && r1 + 1 < nrows;
        const R *x1 = x + r1 * N + tl;
        float2 v[E];
#pragma unroll
        for (int i = 0; i < E; ++i)
            v[i] = make_float2(ld_real(x1 + TPT * i), two ? ld_real(x1 + N + TPT * i) : 0.0f);
        block_fft<LOG2N, -1, kLdsPad, false, kTwPowers>(v, lds + w * SLICE, tl, tw, true);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const float2 hk = spec[tl + TPT * i];
            v[i] = CORR ? cmulc(v[i], hk) : cmul(v[i], hk);
        }
        block_fft<LOG2N, +1, kLdsPad, false, kTwPowers>(v, lds + w * SLICE, tl, tw, true);
        if (active) {
            R *y1 = y + r1 * N + tl;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                st_real(y1 + TPT * i, v[i].x * scale);
                if (two) st_real(y1 + N + TPT * i, v[i].y * scale);
            }
        }
    }
}

// sum |h_t| over k filter taps by one 256-thread workgroup in a fixed summation order
// (the same value on every call); valid in thread 0.  It bounds the L1 norm of the
// folded filter.
__device__ __forceinline__ float filter_l1(const float *__restrict__ h, long k, float *part) {
    float acc = 0.0f;
    for (long i = threadIdx.x; i < k; i += 256) acc += fabsf(h[i]);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    return part[0];
}

// l1->y = sum |h_t| over the k filter taps
__global__ void __launch_bounds__(256) k_filter_l1(const float *__restrict__ h, long k, float2 *__restrict__ l1) {
    __shared__ float part[256];
    const float v = filter_l1(h, k, part);
    if (threadIdx.x == 0) l1->y = v;
}

// spectrum /= L, L = l1->y (divided by |K^[0]| when the spectrum was normalised by the
// kernel sum, dc); thread 0 records L in l1->x (the output scale factor)
__global__ void k_spec_l1(float2 *__restrict__ spec, long n, float2 *__restrict__ l1, const float2 *__restrict__ dc,
                          int normalized) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float L = l1->y;
    if (normalized) L /= fabsf(dc->x);
    if (!(L > 0.0f)) L = 1.0f;
    if (i < n) {
        const float s = 1.0f / L;
        spec[i] = make_float2(spec[i].x * s, spec[i].y * s);
    }
    if (i == 0) l1->x = L;
}

__global__ void k_conv_scales(const float *__restrict__ s_in, float *__restrict__ s_out, const float2 *__restrict__ l1,
                              long rows) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) s_out[i] = s_in[i] * l1->x;
}

hipError_t launch_conv_fp16_filter(const float *h, long k, float2 *spec, long n, const float2 *dc, bool normalized,
                                   hipStream_t st) {
    hipLaunchKernelGGL(k_filter_l1, dim3(1), dim3(256), 0, st, h, k, spec + n);
    hipLaunchKernelGGL(k_spec_l1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spec, n, spec + n, dc,
                       normalized ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_conv_scales(const float *s_in, float *s_out, const float2 *l1, long rows, hipStream_t st) {
    hipLaunchKernelGGL(k_conv_scales, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, s_in, s_out, l1, rows);
    return hipGetLastError();
}

// The same steps for every member of a request group, with the single request's
// arithmetic: member q's (sum |h|, L) pair in g->l1[q] (workgroup q), written before
// the spectra, whose last launch stores them divided by L (1.0f / L, the factor
// k_spec_l1 applies); then the output row scales (workgroup row y = member).
__global__ void __launch_bounds__(256) k_filter_l1_group(ConvGroup *__restrict__ g) {
    const int q = blockIdx.x;
    __shared__ float part[256];
    const float v = filter_l1(g->h[q], g->k[q], part);
    if (threadIdx.x == 0) g->l1[q] = make_float2(v > 0.0f ? v : 1.0f, v);
}

// output row scales of every member, s_out = s_in * L: L from the descriptor (inline
// filters, written by k_filter_l1_group) or from the value after the member's n spectrum
// values (registered filters, written by k_spec_l1 when the filter was registered)
__global__ void k_conv_scales_group(const ConvGroup *__restrict__ g, long n, int registered) {
    const int q = blockIdx.y;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g->real_rows[q]) {
        const float L = registered ? g->spec[q][n].x : g->l1[q].x;
        g->yscale[q][i] = g->xscale[q][i] * L;
    }
}

hipError_t launch_conv_fp16_filter_group(ConvGroup *g, int members, hipStream_t st) {
    if (members < 1 || members > kGroupMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_filter_l1_group, dim3((unsigned)members), dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t launch_conv_scales_group(const ConvGroup *g, int members, long max_rows, long n, bool registered,
                                    hipStream_t st) {
    if (members < 1 || members > kGroupMax || max_rows < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_conv_scales_group, dim3((unsigned)((max_rows + 255) / 256), (unsigned)members), dim3(256), 0,
                       st, g, n, registered ? 1 : 0);
    return hipGetLastError();
}

// One real row of n values convolved through complex transforms of M = n/2 values:
// the row read as z[m] = x[2m] + i x[2m+1] has Z = DFT_M(z), and with k' = M - k
// (indices of Z mod M) and W = exp(-2 pi i / n), for k in [0, M]:
//   2 X[k] = (Z[k] + conj Z[k']) - i W^k (Z[k] - conj Z[k'])     (the row's DFT)
//   2 Y[k] = 2 X[k] H[k]                    (H[k] conjugated for a correlation)
//   4 V[k] = (2Y[k] + conj 2Y[k']) + i W^-k (2Y[k] - conj 2Y[k'])
// and the inverse DFT_M of V (1/M) is y[2m] + i y[2m+1].  One thread per pair
// {k, k'}, k in [0, M/2]; at k' the same terms appear conjugated (2 X[k'] = conj E -
// i conj(W^k D) with E, D the sum and difference above, likewise for V).  V overwrites
// Z in place, times 1/(2n), so the unnormalised inverse transform yields y.  In a
// request group (workgroup row y = member q) row q of zv is member q's, with its
// spectrum.
__global__ void __launch_bounds__(256) k_conv_half_merge(float2 *__restrict__ zv, const float2 *__restrict__ spec,
                                                         long n, int corr, const ConvGroup *__restrict__ group) {
    const long M = n >> 1;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > (M >> 1)) return;
    if (group) {
        zv += (long)blockIdx.y * M;
        spec = group->spec[blockIdx.y];
    }
    const long kp = M - k;  // in [M/2, M]; Z[M] is Z[0]
    const float2 zk = zv[k], zp = zv[kp == M ? 0 : kp];
    double sd, cd;
    sincospi(2.0 * (double)k / (double)n, &sd, &cd);
    const float2 w = make_float2((float)cd, -(float)sd);  // W^k
    const float2 e = make_float2(zk.x + zp.x, zk.y - zp.y);  // Z[k] + conj Z[k']
    const float2 d = make_float2(zk.x - zp.x, zk.y + zp.y);  // Z[k] - conj Z[k']
    const float2 a = cmul(d, w);                               // W^k D
    // 2X[k] = E - i A, 2X[k'] = conj E - i conj A  (-i z = (z.y, -z.x))
    const float2 xk = make_float2(e.x + a.y, e.y - a.x);
    const float2 xp = make_float2(e.x - a.y, -e.y - a.x);
    const float2 hk = spec[k], hp = spec[kp];
    const float2 yk = corr ? cmulc(xk, hk) : cmul(xk, hk);
    const float2 yp = corr ? cmulc(xp, hp) : cmul(xp, hp);
    const float2 f = make_float2(yk.x + yp.x, yk.y - yp.y);  // 2Y[k] + conj 2Y[k']
    const float2 g = make_float2(yk.x - yp.x, yk.y + yp.y);  // 2Y[k] - conj 2Y[k']
    const float2 b = cmulc(g, w);                              // W^-k G
    // 4V[k] = F + i B, 4V[k'] = conj F + i conj B  (i z = (-z.y, z.x))
    const float sc = 0.5f / (float)n;
    zv[k] = make_float2((f.x - b.y) * sc, (f.y + b.x) * sc);
    if (kp < M && kp != k) zv[kp] = make_float2((f.x + b.y) * sc, (-f.y + b.x) * sc);
}

hipError_t launch_conv_half_merge(float2 *zv, const float2 *spec, long n, bool corr, hipStream_t st,
                                  const ConvGroup *group, int members) {
    if (n < 4 || (n & (n - 1))) return hipErrorInvalidValue;
    if (group ? (members < 1 || members > kGroupMax) : !spec) return hipErrorInvalidValue;
    const long threads = (n >> 2) + 1;  // k in [0, n/4]
    hipLaunchKernelGGL(k_conv_half_merge, dim3((unsigned)((threads + 255) / 256), group ? (unsigned)members : 1u),
                       dim3(256), 0, st, zv, spec, n, corr ? 1 : 0, group);
    return hipGetLastError();
}

template <class K>
static long conv_resident(K kernel, int threads) {
    return co_resident_blocks(reinterpret_cast<const void *>(kernel), threads);
}

// Overlap-save circular convolution of rows of n > L values with a filter of K <= L taps
// through L-point transforms (one launch, one HBM read of each window and one write of
// each output).  Window w of the batch (row r = w / W, block b = w mod W) holds the L
// input values x_r[(b M - pre + i) mod n], i < L; its circular L-point convolution with
// the filter equals the circular n-point one for i in [pre, pre + M) whenever
// pre >= K - 1 (a convolution; any pre >= 0 for a correlation, whose filter reaches
// forward) and pre + M <= L - K + 1 + (K - 1 for a convolution, 0 for a correlation).
// Those outputs are positions p = b M + i - pre of the row's result, stored at
// (p + oshift) mod n for p < n (oshift undoes the filter rotation of ESTL_CONV_CENTER,
// which rotates the output instead).  Two windows (any two, of any rows) share one
// complex transform, re and im; the spectrum is the filter's L-point spectrum.
struct OlsArgs {
    const float *s_in;  // FP16 storage: per-row input scales, and the output scales the
    float *s_out;       // launch writes (s_in x L); null otherwise
    long rows;    // real rows of n values
    long W;       // windows per row, ceil(n / M)
    long oshift;  // output rotation, in [0, n)
    int log2n;
    int M;        // valid outputs per window
    int pre;      // first valid position in a window
};

// Four waves per SIMD (at most 128 VGPRs): two 512-thread workgroups per CU at L = 2^13,
// one of 1024 threads at 2^14.
template <int LOG2L, bool CORR, typename R>
__global__ void __launch_bounds__(BlockGeom<LOG2L>::TPT) __attribute__((amdgpu_waves_per_eu(4)))
k_conv_ols(const R *__restrict__ x, R *__restrict__ y, OlsArgs a, const float2 *__restrict__ spec,
           const float2 *__restrict__ tw, float scale) {
    using G = BlockGeom<LOG2L>;
    constexpr int E = G::E, TPT = G::TPT;
    __shared__ float2 lds[Lds<kLdsPad>::size(G::N)];
    if (a.s_out && blockIdx.x == 0) {
        // FP16 storage: the output row scales, s_in times L (the value after the
        // spectrum's L bins), as k_conv_scales computes them
        const float L = spec[(size_t)1 << LOG2L].x;
        for (long r = threadIdx.x; r < a.rows; r += TPT) a.s_out[r] = a.s_in[r] * L;
    }
    const long windows = a.rows * a.W;
    const long pairs = (windows + 1) / 2;
    // offsets inside a row in 32 bits (n <= 2^30): the row bases are workgroup-uniform,
    // so loads and stores take a scalar base and a 32-bit lane offset; x[(s + i) mod n]
    // with s >= -pre wraps correctly in unsigned arithmetic because n divides 2^32
    const unsigned nm = (unsigned)((1L << a.log2n) - 1);
    for (long pr = blockIdx.x; pr < pairs; pr += gridDim.x) {  // uniform across the workgroup
        const int tl = opaque_lane(threadIdx.x);
        const long w0 = 2 * pr, w1 = w0 + 1;
        const bool two = w1 < windows;
        // rows and windows, held in scalar registers (rows < 2^31, windows per row < 2^17)
        const int r0 = __builtin_amdgcn_readfirstlane((int)(w0 / a.W));
        const int b0 = __builtin_amdgcn_readfirstlane((int)(w0 - (long)r0 * a.W));
        const int r1 = two ? __builtin_amdgcn_readfirstlane((int)(w1 / a.W)) : r0;
        const int b1 = two ? __builtin_amdgcn_readfirstlane((int)(w1 - (long)r1 * a.W)) : b0;
        const R *x0 = x + ((long)r0 << a.log2n), *x1 = x + ((long)r1 << a.log2n);
        const unsigned m0 = (unsigned)b0 * (unsigned)a.M, m1 = (unsigned)b1 * (unsigned)a.M;  // first outputs
        const unsigned s0 = m0 - (unsigned)a.pre + (unsigned)tl, s1 = m1 - (unsigned)a.pre + (unsigned)tl;
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            v[e] = make_float2(ld_real(x0 + ((s0 + (unsigned)(TPT * e)) & nm)),
                               two ? ld_real(x1 + ((s1 + (unsigned)(TPT * e)) & nm)) : 0.0f);
        }
        block_fft<LOG2L, -1, kLdsPad, false, kTwPowers>(v, lds, tl, tw, true);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float2 hk = spec[tl + TPT * e];
            v[e] = CORR ? cmulc(v[e], hk) : cmul(v[e], hk);
        }
        block_fft<LOG2L, +1, kLdsPad, false, kTwPowers>(v, lds, tl, tw, true);
        const unsigned n = nm + 1u, osh = (unsigned)a.oshift;
        R *y0 = y + ((long)r0 << a.log2n), *y1 = y + ((long)r1 << a.log2n);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int q = tl + TPT * e - a.pre;  // position in the window's valid outputs
            if (q < 0 || q >= a.M) continue;
            const unsigned p0 = m0 + (unsigned)q, p1 = m1 + (unsigned)q;
            if (p0 < n) st_real(y0 + ((p0 + osh) & nm), v[e].x * scale);
            if (two && p1 < n) st_real(y1 + ((p1 + osh) & nm), v[e].y * scale);
        }
    }
}

template <int LOG2L, bool CORR, typename R>
static hipError_t launch_conv_ols_t(const void *x, void *y, const OlsArgs &a, const float2 *spec, const float2 *tw,
                                    hipStream_t st) {
    using G = BlockGeom<LOG2L>;
    auto kernel = k_conv_ols<LOG2L, CORR, R>;
    const long resident = conv_resident(kernel, G::TPT);
    const long pairs = (a.rows * a.W + 1) / 2;
    long blocks = pairs;
    if (blocks > 4 * resident) blocks = 4 * resident;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(G::TPT), 0, st, static_cast<const R *>(x),
                       static_cast<R *>(y), a, spec, tw, 1.0f / (float)G::N);
    return hipGetLastError();
}

template <int LOG2L>
static hipError_t launch_conv_ols_l(bool fp16, bool corr, const void *x, void *y, long rows, int log2n, long k,
                                    long oshift, const float2 *spec, const float2 *tw, const float *s_in,
                                    float *s_out, hipStream_t st) {
    const long n = 1L << log2n;
    // windows on 128-byte boundaries (ols_first_valid, ols_outputs): every window's loads,
    // and its stores when unrotated, are whole cache lines, for at most 63 of the
    // L - K + 1 outputs a window could give
    const long pre_a = ols_first_valid(k, corr, fp16);
    const long M = ols_outputs(LOG2L, k, corr, fp16);
    if (log2n <= LOG2L || k < 1 || M < 1 || rows < 1) return hipErrorInvalidValue;
    if (fp16 != (s_out != nullptr) || (s_out && !s_in)) return hipErrorInvalidValue;
    OlsArgs a;
    a.s_in = s_in;
    a.s_out = s_out;
    a.rows = rows;
    a.W = (n + M - 1) / M;
    a.oshift = ((oshift % n) + n) % n;
    a.log2n = log2n;
    a.M = (int)M;
    a.pre = (int)pre_a;
    if (fp16)
        return corr ? launch_conv_ols_t<LOG2L, true, _Float16>(x, y, a, spec, tw, st)
                    : launch_conv_ols_t<LOG2L, false, _Float16>(x, y, a, spec, tw, st);
    return corr ? launch_conv_ols_t<LOG2L, true, float>(x, y, a, spec, tw, st)
                : launch_conv_ols_t<LOG2L, false, float>(x, y, a, spec, tw, st);
}

// x, y: rows of n FP32 (fp16 = false) or binary16 reals; spectrum and power table
// (block_fft kTwPowers) of L = 2^log2L, log2L = 13 (512-thread workgroups, two per CU)
// or 14 (1024 threads, one).  FP16 storage: s_in / s_out the per-row scales (the
// launch also writes s_out = s_in x L, L stored after the spectrum); null for FP32.
hipError_t launch_conv_ols(bool fp16, bool corr, int log2L, const void *x, void *y, long rows, int log2n, long k,
                           long oshift, const float2 *spec, const float2 *tw, const float *s_in, float *s_out,
                           hipStream_t st) {
    if (log2L == kOlsLog2)
        return launch_conv_ols_l<kOlsLog2>(fp16, corr, x, y, rows, log2n, k, oshift, spec, tw, s_in, s_out, st);
    if (log2L == kOlsWideLog2)
        return launch_conv_ols_l<kOlsWideLog2>(fp16, corr, x, y, rows, log2n, k, oshift, spec, tw, s_in, s_out, st);
    return hipErrorInvalidValue;
}

template <int LOG2N, bool CORR>
static hipError_t launch_conv_rows_t(const float *x, float *y, long rows, const float2 *spec, const float2 *tw,
                                     float scale, hipStream_t st, const ConvGroup *group) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    auto kernel = k_conv_rows<LOG2N, TPW, CORR, float>;
    const long resident = conv_resident(kernel, G::TPT * TPW);
    const long pairs = (rows + 1) / 2;
    long blocks = (pairs + TPW - 1) / TPW;
    if (blocks > 4 * resident) blocks = 4 * resident;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(G::TPT * TPW), 0, st, x, y, rows, spec, tw, scale, group);
    return hipGetLastError();
}

template <int LOG2N, bool CORR>
static hipError_t launch_conv_rows_h_t(const void *x, void *y, long rows, const float2 *spec, const float2 *tw,
                                       float scale, hipStream_t st, const ConvGroup *group) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    auto kernel = k_conv_rows<LOG2N, TPW, CORR, _Float16>;
    const long resident = conv_resident(kernel, G::TPT * TPW);
    const long pairs = (rows + 1) / 2;
    long blocks = (pairs + TPW - 1) / TPW;
    if (blocks > 4 * resident) blocks = 4 * resident;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(G::TPT * TPW), 0, st, static_cast<const _Float16 *>(x),
                       static_cast<_Float16 *>(y), rows, spec, tw, scale, group);
    return hipGetLastError();
}

template <bool CORR>
static hipError_t launch_conv_rows_h_c(int log2n, const void *x, void *y, long rows, const float2 *spec,
                                       const float2 *tw, float scale, hipStream_t st, const ConvGroup *group) {
    switch (log2n) {
#define CONV_H_CASE(L) \
    case L: return launch_conv_rows_h_t<L, CORR>(x, y, rows, spec, tw, scale, st, group);
        CONV_H_CASE(1) CONV_H_CASE(2) CONV_H_CASE(3) CONV_H_CASE(4) CONV_H_CASE(5) CONV_H_CASE(6) CONV_H_CASE(7)
        CONV_H_CASE(8) CONV_H_CASE(9) CONV_H_CASE(10) CONV_H_CASE(11) CONV_H_CASE(12) CONV_H_CASE(13)
#undef CONV_H_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_conv_rows_fp16(int log2n, bool corr, const void *x, void *y, long rows, const float2 *spec,
                                 const float2 *tw, float scale, hipStream_t st, const ConvGroup *group) {
    return corr ? launch_conv_rows_h_c<true>(log2n, x, y, rows, spec, tw, scale, st, group)
                : launch_conv_rows_h_c<false>(log2n, x, y, rows, spec, tw, scale, st, group);
}

template <bool CORR>
static hipError_t launch_conv_rows_c(int log2n, const float *x, float *y, long rows, const float2 *spec,
                                     const float2 *tw, float scale, hipStream_t st, const ConvGroup *group) {
    switch (log2n) {
#define CONV_CASE(L) \
    case L: return launch_conv_rows_t<L, CORR>(x, y, rows, spec, tw, scale, st, group);
        CONV_CASE(1) CONV_CASE(2) CONV_CASE(3) CONV_CASE(4) CONV_CASE(5) CONV_CASE(6) CONV_CASE(7) CONV_CASE(8)
        CONV_CASE(9) CONV_CASE(10) CONV_CASE(11) CONV_CASE(12) CONV_CASE(13)
#undef CONV_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_conv_rows(int log2n, bool corr, const float *x, float *y, long rows, const float2 *spec,
                            const float2 *tw, float scale, hipStream_t st, const ConvGroup *group) {
    return corr ? launch_conv_rows_c<true>(log2n, x, y, rows, spec, tw, scale, st, group)
                : launch_conv_rows_c<false>(log2n, x, y, rows, spec, tw, scale, st, group);
}

}  // namespace estl
