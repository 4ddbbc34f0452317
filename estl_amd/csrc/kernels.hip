// kernels.hip -- gfx950 kernels of the FFT hot path and their launchers.
//
//   k_fft_rows     batched single-launch FFT of contiguous rows, N <= 2^13: one HBM read
//                  and one HBM write per element (replaces the ceil(log2N/4) launches,
//                  and HBM round trips, of FFT_OpenCL_Contiguous::execute,
//                  FFT_OpenCL_Contiguous.h:108-124)
//   (k_fft_pass, the global Stockham pass of multi-pass plans, is in pass.hip)
//   k_twiddle      twiddle tables evaluated in FP64 on the GPU and rounded to FP32
//                  (replaces the per-element native_cos/native_sin of
//                  KernelGenerator.h:48)
//   k_fill_uniform seeded synthetic input (splitmix64 stream)
//   k_stream_copy  16-byte streaming copy (the achievable-bandwidth reference)
#include <map>
#include <mutex>
#include <utility>

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

// ------------------------------------------------------------------ streaming copy
// The plainest HBM-bound kernel: each thread moves U 16-byte blocks (a wave's loads and
// stores are 1 KiB contiguous each), one pass, no reuse, optionally nontemporal (as the
// transforms' streamed rows are).  Its rate is what a read-once / write-once kernel
// reaches on the device.  U = 1 measured fastest on MI355X (512 MiB: 6.4 TB/s
// nontemporal, against 6.1 at U = 4 and 4.3 at U = 8; 8-byte accesses 6.2 TB/s).
typedef unsigned v4u_t __attribute__((ext_vector_type(4)));
constexpr int kStreamCopyU = 1;
template <bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const v4u_t *__restrict__ src, v4u_t *__restrict__ dst, long n16) {
    constexpr int U = kStreamCopyU;
    const long base = (long)blockIdx.x * (256 * U) + threadIdx.x;
    v4u_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const long i = base + 256L * u;
        if (i < n16) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const long i = base + 256L * u;
        if (i < n16) {
            if (NT)
                __builtin_nontemporal_store(v[u], dst + i);
            else
                dst[i] = v[u];
        }
    }
}

hipError_t launch_stream_copy(const void *src, void *dst, long n16, bool nontemporal, hipStream_t st) {
    if (n16 <= 0) return hipSuccess;
    const long per_block = 256L * kStreamCopyU;
    const long blocks = (n16 + per_block - 1) / per_block;
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    if (nontemporal)
        hipLaunchKernelGGL(k_stream_copy<true>, dim3((unsigned)blocks), dim3(256), 0, st,
                           static_cast<const v4u_t *>(src), static_cast<v4u_t *>(dst), n16);
    else
        hipLaunchKernelGGL(k_stream_copy<false>, dim3((unsigned)blocks), dim3(256), 0, st,
                           static_cast<const v4u_t *>(src), static_cast<v4u_t *>(dst), n16);
    return hipGetLastError();
}

// ------------------------------------------------------------------ batched rows
// Rows of N <= 2^13 points: a group of TPW = max(1, 256 / TPT) rows, one per TPT = N/16
// threads (16 values per thread), per workgroup, and one workgroup per group -- the
// whole transform in registers and padded LDS, one HBM read and one HBM write per
// value, nontemporal loads and stores.  A partial last group computes its missing rows
// on a copy of the last row and stores nothing for them.  With twiddles by powers
// (kRowsTw) the kernel holds 68 VGPRs at 4096 points and 75 at 8192, so LDS alone
// bounds residency (four 34 KiB groups per CU; two 8192-point rows).  Measured against
// the earlier persistent grid (four resident rounds of workgroups, the next group
// prefetched into registers; at 4096 and 8192 points the block table's twiddles held in
// registers across rows, 167 / 206 VGPRs): 6.5 % faster at 4096 x 16384 and 9 % at
// 8192 x 8192, 3-6 % at 512-2048 points -- many short-lived workgroups keep more loads
// in flight, and the grid's last round is one group deep.
template <int LOG2N, int DIR, int TPW, bool SCALED>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT * TPW)
k_fft_rows(const float2 *__restrict__ in, float2 *__restrict__ out, long batch,
           const float2 *__restrict__ tw, float scale) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    constexpr int SLICE = G::uses_lds ? Lds<kLdsPad>::size(N) : 0;
    constexpr int LDS = G::uses_lds ? TPW * SLICE : 1;
    __shared__ float2 lds[LDS];
    float2 v[E];
    if constexpr (TPW == 1) {
        // one row per workgroup, the grid one workgroup per row
        const int tl = threadIdx.x;
        const long row = blockIdx.x;
        const float2 *src = in + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = ld_stream(src + TPT * i);
        block_fft<LOG2N, DIR, kLdsPad, false, kRowsTw<LOG2N>>(v, lds, tl, tw, false);
        float2 *dst = out + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) st_stream(dst + TPT * i, SCALED ? cscale(v[i], scale) : v[i]);
    } else {
        const int t = threadIdx.x;
        const int w = t / TPT, tl = t % TPT;
        const long want = (long)blockIdx.x * TPW + w;
        const long row = want < batch ? want : batch - 1;
        const float2 *src = in + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = ld_stream(src + TPT * i);
        {
            const int to = opaque_lane(t);
            block_fft<LOG2N, DIR, kLdsPad, false, kRowsTw<LOG2N>>(v, lds + (to / TPT) * SLICE, to % TPT, tw, false);
        }
        if (want < batch) {
            float2 *dst = out + row * N + tl;
#pragma unroll
            for (int i = 0; i < E; ++i) st_stream(dst + TPT * i, SCALED ? cscale(v[i], scale) : v[i]);
        }
    }
}

// Rows on the reference's tile-planarized layout (PlannarizedComplexArray<5>): the same
// transform, loads and stores by element index through planar_pos.  The workgroups walk
// the groups of TPW rows with a stride of the grid, which the launcher sizes to one
// workgroup per group.
template <int LOG2N, int DIR, int TPW>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT * TPW)
k_fft_rows_planar(const float2 *__restrict__ in, float2 *__restrict__ out, long batch, const float2 *__restrict__ tw,
                  float scale) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    constexpr int SLICE = G::uses_lds ? Lds<kLdsPad>::size(N) : 0;
    constexpr int LDS = G::uses_lds ? TPW * SLICE : 1;
    __shared__ float2 lds[LDS];
    const long groups = (batch + TPW - 1) / TPW;
    bool busy = false;
    for (long grp = blockIdx.x; grp < groups; grp += gridDim.x) {  // uniform across the workgroup
        const int t = opaque_lane(threadIdx.x);
        const int w = t / TPT, tl = t % TPT;
        const long want = grp * TPW + w;
        const long row = want < batch ? want : batch - 1;
        float2 v[E];
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = ld_planar(in, row * N + tl + TPT * i);
        block_fft<LOG2N, DIR, kLdsPad, false, kRowsTw<LOG2N>>(v, lds + w * SLICE, tl, tw, busy);
        busy = true;
        if (want < batch) {
#pragma unroll
            for (int i = 0; i < E; ++i) st_planar(out, row * N + tl + TPT * i, cscale(v[i], scale));
        }
    }
}

// ------------------------------------------------------------------ launchers
// Workgroups of `kernel` resident on the whole device at once (occupancy x CUs),
// cached per kernel (by address: kernels of one signature share a pointer type) and
// device, under a lock (launches come from several threads and devices).
long co_resident_blocks(const void *kernel, int block_threads) {
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, long> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const auto key = std::make_pair(kernel, dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return cache[key] = (long)per_cu * cus;
}

template <int LOG2N, int DIR>
static hipError_t launch_rows_t(const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                                hipStream_t st) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    const long groups = (batch + TPW - 1) / TPW;  // one workgroup each
    if (groups > 0x7fffffffL) return hipErrorInvalidValue;
    auto kernel = (scale != 1.0f) ? k_fft_rows<LOG2N, DIR, TPW, true> : k_fft_rows<LOG2N, DIR, TPW, false>;
    hipLaunchKernelGGL(kernel, dim3((unsigned)groups), dim3(G::TPT * TPW), 0, st, in, out, batch, tw, scale);
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

template <int LOG2N, int DIR>
static hipError_t launch_rows_planar_t(const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                                       hipStream_t st) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    auto kernel = k_fft_rows_planar<LOG2N, DIR, TPW>;
    // one workgroup per group of rows (2-3.5 % faster than four resident rounds at 4096
    // and 8192 points)
    const long groups = (batch + TPW - 1) / TPW;
    if (groups > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, dim3((unsigned)groups), dim3(G::TPT * TPW), 0, st, in, out, batch, tw, scale);
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_rows_planar_d(int log2n, const float2 *in, float2 *out, long batch, const float2 *tw,
                                       float scale, hipStream_t st) {
    switch (log2n) {
#define ROWS_P_CASE(L) \
    case L: return launch_rows_planar_t<L, DIR>(in, out, batch, tw, scale, st);
        ROWS_P_CASE(1) ROWS_P_CASE(2) ROWS_P_CASE(3) ROWS_P_CASE(4) ROWS_P_CASE(5) ROWS_P_CASE(6) ROWS_P_CASE(7)
        ROWS_P_CASE(8) ROWS_P_CASE(9) ROWS_P_CASE(10) ROWS_P_CASE(11) ROWS_P_CASE(12) ROWS_P_CASE(13)
#undef ROWS_P_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows_planar(int log2n, int dir, const float2 *in, float2 *out, long batch, const float2 *tw,
                              float scale, hipStream_t st) {
    return dir < 0 ? launch_rows_planar_d<-1>(log2n, in, out, batch, tw, scale, st)
                   : launch_rows_planar_d<+1>(log2n, in, out, batch, tw, scale, st);
}

long rows_twiddle_count(int log2n) {
    switch (log2n) {
#define TW_CASE(L) \
    case L: return BlockGeom<L>::twsize;
        TW_CASE(1) TW_CASE(2) TW_CASE(3) TW_CASE(4) TW_CASE(5) TW_CASE(6) TW_CASE(7) TW_CASE(8)
        TW_CASE(9) TW_CASE(10) TW_CASE(11) TW_CASE(12) TW_CASE(13) TW_CASE(14)
#undef TW_CASE
        default: return -1;
    }
}

long rows_power_count(int log2n) {
    switch (log2n) {
#define POW_CASE(L) \
    case L: return BlockGeom<L>::powsize;
        POW_CASE(1) POW_CASE(2) POW_CASE(3) POW_CASE(4) POW_CASE(5) POW_CASE(6) POW_CASE(7) POW_CASE(8)
        POW_CASE(9) POW_CASE(10) POW_CASE(11) POW_CASE(12) POW_CASE(13) POW_CASE(14)
#undef POW_CASE
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
