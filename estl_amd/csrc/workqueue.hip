// workqueue.hip -- k_ubatch: a batch of transforms of uneven power-of-two sizes in one
// persistent launch driven by a work queue.
//
// The reference batches only transforms of one size (batchCount, FFT.h:25-27) and
// serves mixed sizes with one FFT object -- one plan, one program, one launch per pass --
// per size (tests/Test.cpp:48-53).  Here a descriptor list holds every transform of the
// batch and one launch computes them all:
//   * a work item is either a group of same-size rows (n <= 2^12: kUThreads/TPT rows computed
//     like k_fft_rows) or one whole large transform (2^13 <= n <= 2^20: both passes of a
//     two-pass plan, tile by tile, through a scratch region, with a workgroup barrier
//     between the passes -- no dependency crosses workgroups);
//   * items are ordered largest first and handed out by an atomic ticket, so workgroups
//     that drew small items keep drawing while others finish large ones;
//   * the last workgroup to leave resets the ticket counters, so the launch can be
//     replayed (and captured in a HIP graph) without a memset.
#include <utility>

#include "pass_tile.h"
#include "ubatch.h"

namespace estl {

// workgroup size 2^kUThreadsLog2: small items take 2^kUThreadsLog2 / TPT rows, large
// items tiles of 2^(kUThreadsLog2 + 4 - B) columns.  Measured on MI355X (8192 sizes
// 2^U{8..18}): 512 threads at four waves per SIMD 2.53 ms, 256 threads 3.42 ms (its
// 8-column tiles of radix 2^9 read 64-byte runs), 512 threads at three waves 2.97 ms.
constexpr int kUThreads = 1 << kUThreadsLog2;
static_assert(kUThreadsLog2 + 4 <= kUMaxSmallLog2 + 1, "every large size must fill a workgroup's tile");

// pass split of the large sizes: 2^L = 2^(L/2) x 2^(L - L/2).  Measured on MI355X
// (1 GiB of one size per launch): the even split beats skewed ones -- a radix-16 first
// pass with 256-column tiles and a 2^10..2^12 second pass was 1.7x (L = 14) to 6x
// (L = 16) slower, a 2^9 second pass 2-20 % slower.
__host__ __device__ constexpr int ubatch_split(int L) { return L / 2; }

template <int L>
struct USmall {
    using G = BlockGeom<L>;
    static constexpr int TPW = kUThreads / G::TPT;
    static constexpr int slice = G::uses_lds ? Lds<kLdsPad>::size(G::N) : 0;
    static constexpr int lds = TPW * slice;
};
template <int L>
struct ULarge {
    static constexpr int B1 = ubatch_split(L), B2 = L - B1;
    static constexpr int T1 = kUThreadsLog2 + 4 - B1, T2 = kUThreadsLog2 + 4 - B2;  // threads = T B / 16
    static constexpr int lds = PassTile<B1, T1>::lds_size > PassTile<B2, T2>::lds_size ? PassTile<B1, T1>::lds_size
                                                                                        : PassTile<B2, T2>::lds_size;
};

template <int... Ls>
constexpr int small_lds_max(std::integer_sequence<int, Ls...>) {
    int m = 0;
    for (int v : {USmall<Ls + 1>::lds...}) m = v > m ? v : m;
    return m;
}
template <int... Ls>
constexpr int large_lds_max(std::integer_sequence<int, Ls...>) {
    int m = 0;
    for (int v : {ULarge<Ls + kUMaxSmallLog2 + 1>::lds...}) m = v > m ? v : m;
    return m;
}
// LDS of the kernel: the largest need over every size class it dispatches
constexpr int kULdsSmall = small_lds_max(std::make_integer_sequence<int, kUMaxSmallLog2>{});
constexpr int kULdsLarge = large_lds_max(std::make_integer_sequence<int, kUMaxLog2 - kUMaxSmallLog2>{});
constexpr int kULds = kULdsSmall > kULdsLarge ? kULdsSmall : kULdsLarge;
static_assert(kULds * 8 <= (kUThreadsLog2 >= 10 ? 160 : 80) * 1024, "work-queue LDS over budget");

template <int L, int DIR>
__device__ __forceinline__ void small_item(const UArgs &a, const UItem &it, float2 *lds) {
    using G = BlockGeom<L>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT, TPW = USmall<L>::TPW;
    const int t = opaque_lane(threadIdx.x), w = t / TPT, tl = t % TPT;
    const bool active = w < it.count;
    const long off = a.offs[it.first + (active ? w : 0)];
    float2 v[E];
    const float2 *src = a.in + off + tl;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = ld_stream(src + TPT * i);
    block_fft<L, DIR, kLdsPad, false, kTwPowers>(v, lds + w * USmall<L>::slice, tl, a.tw.pow[L], true);
    if (active) {
        float2 *dst = a.out + off + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) st_stream(dst + TPT * i, cscale(v[i], a.scale));
    }
    (void)N;
    (void)TPW;
}

template <int L, int DIR>
__device__ __forceinline__ void large_item(const UArgs &a, const UItem &it, float2 *lds) {
    constexpr int B1 = ULarge<L>::B1, B2 = ULarge<L>::B2, T1 = ULarge<L>::T1, T2 = ULarge<L>::T2;
    const long off = a.offs[it.first];
    PassArgs p1{};
    p1.batch = 1;
    p1.A = 1;
    p1.log2n = L;
    p1.log2S = L - B1;
    p1.scale = 1.0f;
    p1.tw_local = a.tw.pow[B1];
#pragma unroll 1
    for (long tile = 0; tile < (1L << (L - B1 - T1)); ++tile)
        pass_tile<B1, DIR, T1>(a.in + off, a.scratch + off, p1, tile, lds, true);
    __syncthreads();  // this workgroup's scratch writes are visible to all its threads
    PassArgs p2{};
    p2.batch = 1;
    p2.A = 1L << B1;
    p2.log2n = L;
    p2.log2S = 0;
    p2.lobits = a.tw.lobits[L];
    p2.scale = a.scale;
    p2.tw_local = a.tw.pow[B2];
    p2.tw_lo = a.tw.lo[L];
    p2.tw_hi = a.tw.hi[L];
#pragma unroll 1
    for (long tile = 0; tile < (1L << (L - B2 - T2)); ++tile)
        pass_tile<B2, DIR, T2>(a.scratch + off, a.out + off, p2, tile, lds, true);
}

// waves per SIMD the register budget is held to (0: the compiler's choice)
#ifndef ESTL_UBATCH_WAVES
#define ESTL_UBATCH_WAVES 4
#endif
template <int DIR>
__global__ void __launch_bounds__(kUThreads)
#if ESTL_UBATCH_WAVES
__attribute__((amdgpu_waves_per_eu(ESTL_UBATCH_WAVES, ESTL_UBATCH_WAVES)))
#endif
k_ubatch(UArgs a) {
    __shared__ float2 lds[kULds];
    __shared__ int s_item;
#pragma unroll 1
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&a.counters[0], 1);
        __syncthreads();
        // the ticket and the descriptor are the same for the whole workgroup: say so, so
        // the size switch below is a scalar branch rather than predicated code whose
        // cases would all be live at once
        const int idx = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (idx >= a.n_items) break;  // every workgroup reaches this exit
        const UItem raw = a.items[idx];
        UItem it;
        it.log2n = __builtin_amdgcn_readfirstlane(raw.log2n);
        it.count = __builtin_amdgcn_readfirstlane(raw.count);
        it.first = __builtin_amdgcn_readfirstlane(raw.first);
        it.pad = 0;
        switch (it.log2n) {
#define SMALL(L) \
    case L: small_item<L, DIR>(a, it, lds); break;
#define LARGE(L) \
    case L: large_item<L, DIR>(a, it, lds); break;
            SMALL(1) SMALL(2) SMALL(3) SMALL(4) SMALL(5) SMALL(6) SMALL(7) SMALL(8) SMALL(9) SMALL(10) SMALL(11)
            SMALL(12)
#if ESTL_UBATCH_TLOG >= 10
            SMALL(13)
#else
            LARGE(13)
#endif
            LARGE(14) LARGE(15) LARGE(16) LARGE(17) LARGE(18) LARGE(19) LARGE(20)
#undef SMALL
#undef LARGE
            default: break;
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&a.counters[1], 1) == (int)gridDim.x - 1) {
            atomicExch(&a.counters[0], 0);
            atomicExch(&a.counters[1], 0);
        }
    }
}

long ubatch_resident_blocks() {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ubatch<-1>, kUThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return (long)per_cu * cus;
}

int ubatch_small_rows_per_item(int log2n) {
    switch (log2n) {
#define TPWC(L) \
    case L: return USmall<L>::TPW;
        TPWC(1) TPWC(2) TPWC(3) TPWC(4) TPWC(5) TPWC(6) TPWC(7) TPWC(8) TPWC(9) TPWC(10) TPWC(11) TPWC(12)
#if ESTL_UBATCH_TLOG >= 10
        TPWC(13)
#endif
#undef TPWC
        default: return 1;
    }
}

hipError_t launch_ubatch(int dir, const UArgs &a, long blocks, hipStream_t st) {
    if (dir < 0)
        hipLaunchKernelGGL(k_ubatch<-1>, dim3((unsigned)blocks), dim3(kUThreads), 0, st, a);
    else
        hipLaunchKernelGGL(k_ubatch<+1>, dim3((unsigned)blocks), dim3(kUThreads), 0, st, a);
    return hipGetLastError();
}

}  // namespace estl
