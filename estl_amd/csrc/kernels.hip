// kernels.hip -- gfx950 kernels of the FFT hot path and their launchers.
//
//   k_fft_rows     batched single-launch FFT of contiguous rows, N <= 2^13: one HBM read
//                  and one HBM write per element (replaces the ceil(log2N/4) launches,
//                  and HBM round trips, of FFT_OpenCL_Contiguous::execute,
//                  FFT_OpenCL_Contiguous.h:108-124)
//   k_fft_pass     one global Stockham pass of radix B (B up to 2^13, itself computed
//                  by block_fft through LDS) for large N: the pass formula of
//                  FFT_OpenCL_Contiguous.h:167-180 with a tile of T adjacent columns per
//                  workgroup so loads and stores stay coalesced
//   k_twiddle      twiddle tables evaluated in FP64 on the GPU and rounded to FP32
//                  (replaces the per-element native_cos/native_sin of
//                  KernelGenerator.h:48)
//   k_fill_uniform seeded synthetic input (splitmix64 stream)
#include "fft_block.h"
#include "internal.h"

namespace estl {

// ------------------------------------------------------------------ twiddles
// mode 0 (per-pass layout of BlockGeom): entry e -> s = e / A + 1, g = e % A, m = g*s
// mode 1 (linear): entry e -> m = e * mult
// value = exp(-2 pi i m / nq), computed as sincospi(2 m / nq) in double.
__global__ void k_twiddle(float2 *__restrict__ dst, long count, int mode, long A, long mult, long nq) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    long m;
    if (mode == 0) {
        long s = e / A + 1, g = e % A;
        m = (g * s) % nq;
    } else {
        m = (e * mult) % nq;
    }
    double sn, cs;
    sincospi(2.0 * (double)m / (double)nq, &sn, &cs);
    dst[e] = make_float2((float)cs, (float)(-sn));
}

// ------------------------------------------------------------------ synthetic data
__global__ void k_fill_uniform(float *__restrict__ dst, long count, uint64_t key, uint64_t offset) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        uint64_t u = splitmix64(key + offset + (uint64_t)i);
        dst[i] = (float)((double)(u >> 40) * (1.0 / 8388608.0) - 1.0);
    }
}

// ------------------------------------------------------------------ batched rows
// A group is TPW rows, one per TPT threads.  The grid is persistent (sized to the
// resident workgroup count by the launcher): each workgroup walks groups
// blockIdx.x, blockIdx.x + gridDim.x, ... and issues the loads of its next group
// before running the current one through LDS, so HBM reads stay in flight during the
// register/LDS passes.  Input and output are touched once, so both use nontemporal
// (streaming) accesses.  Rows past the batch end (a partial last group) are computed
// on whatever the clamped load returned and never stored.
template <int LOG2N, int DIR, int TPW, bool SCALED>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT * TPW)
k_fft_rows(const float2 *__restrict__ in, float2 *__restrict__ out, long batch,
           const float2 *__restrict__ tw, float scale) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    constexpr int SLICE = G::uses_lds ? Lds<kLdsPad>::size(N) : 0;
    constexpr int LDS = G::uses_lds ? TPW * SLICE : 1;
    __shared__ float2 lds[LDS];
    float2 v[E], vn[E];
    bool busy = false;
    if constexpr (TPW == 1) {
        // one row per workgroup and group
        const int tl = threadIdx.x;
        long row = blockIdx.x;
        if (row < batch) {
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = ld_stream(in + row * N + tl + TPT * i);
        }
        for (; row < batch; row += gridDim.x) {
            const long nr = row + gridDim.x;
            if (nr < batch) {
#pragma unroll
                for (int i = 0; i < E; ++i) vn[i] = ld_stream(in + nr * N + tl + TPT * i);
            }
            block_fft<LOG2N, DIR, kLdsPad>(v, lds, tl, tw, busy);
            busy = true;
            float2 *dst = out + row * N + tl;
#pragma unroll
            for (int i = 0; i < E; ++i) st_stream(dst + TPT * i, SCALED ? cscale(v[i], scale) : v[i]);
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = vn[i];
        }
    } else {
        const int t = threadIdx.x;
        const int w = t / TPT, tl = t % TPT;
        const long groups = (batch + TPW - 1) / TPW;
        const long last = batch - 1;
        long grp = blockIdx.x;
        if (grp < groups) {
            const float2 *src = in + min(grp * TPW + w, last) * N + tl;
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = ld_stream(src + TPT * i);
        }
        for (; grp < groups; grp += gridDim.x) {  // uniform across the workgroup
            const long row = grp * TPW + w;
            const long next = grp + gridDim.x;
            if (next < groups) {
                const float2 *src = in + min(next * TPW + w, last) * N + tl;
#pragma unroll
                for (int i = 0; i < E; ++i) vn[i] = ld_stream(src + TPT * i);
            }
            block_fft<LOG2N, DIR, kLdsPad>(v, lds + w * SLICE, tl, tw, busy);
            busy = true;
            if (row < batch) {
                float2 *dst = out + row * N + tl;
#pragma unroll
                for (int i = 0; i < E; ++i) st_stream(dst + TPT * i, SCALED ? cscale(v[i], scale) : v[i]);
            }
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = vn[i];
        }
    }
}

// ------------------------------------------------------------------ global pass
// One pass q of a multi-pass plan over `batch` rows of length n: radix B = 2^LOG2B,
// A = A_q, S = n / (A B).  Column col in [0, batch * n / B) is (row r, j in [0, n/B)),
// g = j / S, z = j % S.  Column c of the tile is handled by threads c + T * tl.
//   load   x_s = in[r n + S (B g + s) + z] * W_{A B}^{g s}   (two-level table)
//   DFT_B  through block_fft (LDS)
//   store  out[r n + j + h n / B] = X_h
// If S >= T the tile's loads are T contiguous values per s (direct, coalesced); else the
// tile's input is one contiguous block of T*B values, staged through LDS.
template <int LOG2B, int DIR, int T>
__global__ void __launch_bounds__(BlockGeom<LOG2B>::TPT * T)
k_fft_pass(const float2 *__restrict__ in, float2 *__restrict__ out, PassArgs a) {
    using G = BlockGeom<LOG2B>;
    constexpr int B = G::N, E = G::E, TPT = G::TPT;
    constexpr int LDSP = Lds<kLdsPad>::size(B);  // one column's block_fft slice (also stages loads)
    __shared__ float2 lds[T * LDSP];
    const int t = threadIdx.x;
    const int c = t % T, tl = t / T;
    // XCD-aware tile order: consecutive tiles land on the same XCD (blocks are dealt
    // round-robin to the 8 XCDs), so tiles that share cache lines share an L2.
    long blk = blockIdx.x;
    if (a.xcd_remap) {
        const long nb = gridDim.x;
        const long per = nb >> 3;
        if ((nb & 7) == 0) blk = (blk & 7) * per + (blk >> 3);
    }
    const long cols_per_row = a.n >> LOG2B;
    const long col0 = blk * T;
    const long r = col0 / cols_per_row;
    const long j0 = col0 - r * cols_per_row;
    const long j = j0 + c;
    const long S = a.S;
    const long g = j / S, z = j - (j / S) * S;
    const float2 *src = in + r * a.n;
    float2 v[E];
    bool lds_busy = false;
    if (S >= T) {
        const float2 *p = src + S * B * g + z;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = p[S * (tl + TPT * i)];
    } else {
        // contiguous block of T*B values starting at B*j0: value o -> column
        // (o / (S B)) * S + o % S, element (o % (S B)) / S
        const float2 *p = src + (long)B * j0;
        constexpr int NT = T * TPT;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int o = t + NT * k;
            const int gl = o / (int(S) * B), rem = o - gl * (int(S) * B);
            const int s = rem / int(S), zz = rem - s * int(S);
            lds[(gl * int(S) + zz) * LDSP + Lds<kLdsPad>::idx(s)] = p[o];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = lds[c * LDSP + Lds<kLdsPad>::idx(tl + TPT * i)];
        lds_busy = true;
    }
    if (a.A > 1) {
        const long lomask = (1L << a.lobits) - 1;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const long m = g * (long)(tl + TPT * i);
            const float2 w = cmul(a.tw_lo[m & lomask], a.tw_hi[m >> a.lobits]);
            v[i] = ctw<DIR>(v[i], w);
        }
    }
    block_fft<LOG2B, DIR, kLdsPad>(v, lds + c * LDSP, tl, a.tw_local, lds_busy);
    float2 *dst = out + r * a.n + j;
    const long stride = cols_per_row;
    if (a.scale != 1.0f) {
#pragma unroll
        for (int i = 0; i < E; ++i) dst[stride * (tl + TPT * i)] = cscale(v[i], a.scale);
    } else {
#pragma unroll
        for (int i = 0; i < E; ++i) dst[stride * (tl + TPT * i)] = v[i];
    }
}

// ------------------------------------------------------------------ launchers
// Workgroups of `kernel` resident on the whole device at once (occupancy x CUs),
// cached per kernel and device.
template <class K>
static long resident_blocks(K kernel, int block_threads) {
    static long cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cache[dev] == 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_threads, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        cache[dev] = (long)per_cu * cus;
    }
    return cache[dev];
}

// persistent grid: this many resident rounds of workgroups
constexpr long kRowsGridRounds = 4;

template <int LOG2N, int DIR>
static hipError_t launch_rows_t(const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                                hipStream_t st) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    const long groups = (batch + TPW - 1) / TPW;
    auto kernel = (scale != 1.0f) ? k_fft_rows<LOG2N, DIR, TPW, true> : k_fft_rows<LOG2N, DIR, TPW, false>;
    long blocks = resident_blocks(k_fft_rows<LOG2N, DIR, TPW, false>, G::TPT * TPW) * kRowsGridRounds;
    if (blocks > groups) blocks = groups;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(G::TPT * TPW), 0, st, in, out, batch, tw, scale);
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_rows_d(int log2n, const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                                hipStream_t st) {
    switch (log2n) {
#define ROWS_CASE(L) \
    case L: return launch_rows_t<L, DIR>(in, out, batch, tw, scale, st);
        ROWS_CASE(1) ROWS_CASE(2) ROWS_CASE(3) ROWS_CASE(4) ROWS_CASE(5) ROWS_CASE(6) ROWS_CASE(7)
        ROWS_CASE(8) ROWS_CASE(9) ROWS_CASE(10) ROWS_CASE(11) ROWS_CASE(12) ROWS_CASE(13)
#undef ROWS_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows(int log2n, int dir, const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                       hipStream_t st) {
    return dir < 0 ? launch_rows_d<-1>(log2n, in, out, batch, tw, scale, st)
                   : launch_rows_d<+1>(log2n, in, out, batch, tw, scale, st);
}

long rows_twiddle_count(int log2n) {
    switch (log2n) {
#define TW_CASE(L) \
    case L: return BlockGeom<L>::twsize;
        TW_CASE(1) TW_CASE(2) TW_CASE(3) TW_CASE(4) TW_CASE(5) TW_CASE(6) TW_CASE(7) TW_CASE(8)
        TW_CASE(9) TW_CASE(10) TW_CASE(11) TW_CASE(12) TW_CASE(13)
#undef TW_CASE
        default: return -1;
    }
}

// per-pass parameters of the block table (for building it): A_q and B_q of pass q
void rows_pass_params(int log2n, int q, long *A, int *B) {
    const int nfull = log2n / 4, rem = log2n % 4;
    long a = 1;
    for (int p = 0; p < q; ++p) a *= (p < nfull) ? 16 : (1 << rem);
    *A = a;
    *B = (q < nfull) ? 16 : (1 << rem);
}
int rows_pass_count(int log2n) { return log2n / 4 + ((log2n % 4) ? 1 : 0); }

int pass_tile(int log2b) {
    // T adjacent columns per workgroup: 16 (128-byte store segments) while the tile
    // fits 64 KiB of LDS values, fewer for the largest radices
    int t = 8192 >> log2b;
    if (t > 16) t = 16;
    if (t < 1) t = 1;
    return t;
}

template <int LOG2B, int DIR>
static hipError_t launch_pass_t(const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    using G = BlockGeom<LOG2B>;
    constexpr int T = (8192 >> LOG2B) > 16 ? 16 : ((8192 >> LOG2B) < 1 ? 1 : (8192 >> LOG2B));
    const long cols = a.batch * (a.n >> LOG2B);
    if (cols % T) return hipErrorInvalidValue;
    const long blocks = cols / T;
    hipLaunchKernelGGL((k_fft_pass<LOG2B, DIR, T>), dim3((unsigned)blocks), dim3(G::TPT * T), 0, st, in, out, a);
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_pass_d(int log2b, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    switch (log2b) {
#define PASS_CASE(L) \
    case L: return launch_pass_t<L, DIR>(in, out, a, st);
        PASS_CASE(4) PASS_CASE(5) PASS_CASE(6) PASS_CASE(7) PASS_CASE(8) PASS_CASE(9) PASS_CASE(10)
        PASS_CASE(11) PASS_CASE(12) PASS_CASE(13)
#undef PASS_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pass(int log2b, int dir, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    return dir < 0 ? launch_pass_d<-1>(log2b, in, out, a, st) : launch_pass_d<+1>(log2b, in, out, a, st);
}

hipError_t launch_twiddle(float2 *dst, long count, int mode, long A, long mult, long nq, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const long blocks = (count + 255) / 256;
    hipLaunchKernelGGL(k_twiddle, dim3((unsigned)blocks), dim3(256), 0, st, dst, count, mode, A, mult, nq);
    return hipGetLastError();
}

hipError_t launch_fill_uniform(float *dst, long count, uint64_t seed, uint64_t offset, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const uint64_t key = [&] {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }();
    long blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)blocks), dim3(256), 0, st, dst, count, key, offset);
    return hipGetLastError();
}

}  // namespace estl
