// pass.hip -- k_fft_pass: one global Stockham pass of a multi-pass plan (large N).
//
// Compiled once per direction (-DESTL_PASS_DIR=-1 / +1) so the two instantiation sets
// build in parallel.
//
// One workgroup per tile of the pass (the tile itself is computed by pass_tile,
// pass_tile.h).
#include "pass_tile.h"

#ifndef ESTL_PASS_DIR
#error "compile pass.hip with -DESTL_PASS_DIR=-1 or +1"
#endif

namespace estl {

// Tile t's number after the optional XCD remap: workgroups are dealt round-robin to the
// 8 XCDs, so consecutive tiles (which share cache lines when tiles are narrower than a
// line) are given to workgroups of one XCD, where they meet in its L2.
__device__ __forceinline__ long remap_tile(long t, long ntiles, int xcd_remap) {
    if (xcd_remap && (ntiles & 7) == 0) return (t & 7) * (ntiles >> 3) + (t >> 3);
    return t;
}

// One tile per workgroup.  (Measured on MI355X: a persistent, software-pipelined form
// of this kernel -- each workgroup loading its next tile while computing the current
// one -- was 10-20 % slower than letting the hardware overlap one-tile workgroups at
// four waves per SIMD.)
template <int LOG2B, int DIR, int LOG2T, bool MID>
__global__ void __launch_bounds__(BlockGeom<LOG2B>::TPT << LOG2T)
k_fft_pass(const float2 *__restrict__ in, float2 *__restrict__ out, PassArgs a) {
    __shared__ float2 lds[PassTile<LOG2B, LOG2T>::lds_size];
    const long tile = remap_tile(blockIdx.x, gridDim.x, a.xcd_remap);
    pass_tile<LOG2B, DIR, LOG2T, MID>(in, out, a, tile, lds, false);
}

template <int LOG2B, int LOG2T, bool MID>
static hipError_t launch_t(const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    if constexpr (LOG2B + LOG2T > kMaxPassTileLog2) {
        return hipErrorInvalidValue;
    } else {
        const int lcpr = a.log2n - LOG2B;
        if (lcpr < LOG2T) return hipErrorInvalidValue;
        const long tiles = a.batch << (lcpr - LOG2T);
        hipLaunchKernelGGL((k_fft_pass<LOG2B, ESTL_PASS_DIR, LOG2T, MID>), dim3((unsigned)tiles),
                           dim3(BlockGeom<LOG2B>::TPT << LOG2T), 0, st, in, out, a);
        return hipGetLastError();
    }
}

template <int LOG2B, bool MID = false>
static hipError_t launch_b(int log2t, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    switch (log2t) {
        case 0: return launch_t<LOG2B, 0, MID>(in, out, a, st);
        case 1: return launch_t<LOG2B, 1, MID>(in, out, a, st);
        case 2: return launch_t<LOG2B, 2, MID>(in, out, a, st);
        case 3: return launch_t<LOG2B, 3, MID>(in, out, a, st);
        case 4: return launch_t<LOG2B, 4, MID>(in, out, a, st);
        case 5: return launch_t<LOG2B, 5, MID>(in, out, a, st);
        case 6: return launch_t<LOG2B, 6, MID>(in, out, a, st);
        default: return hipErrorInvalidValue;
    }
}

#if ESTL_PASS_DIR < 0
#define ESTL_PASS_LAUNCHER launch_pass_fwd
#else
#define ESTL_PASS_LAUNCHER launch_pass_bwd
#endif

hipError_t ESTL_PASS_LAUNCHER(int log2b, int log2t, const float2 *in, float2 *out, const PassArgs &a,
                              hipStream_t st) {
    switch (log2b) {
        case 1: return launch_b<1>(log2t, in, out, a, st);
        case 2: return launch_b<2>(log2t, in, out, a, st);
        case 3: return launch_b<3>(log2t, in, out, a, st);
        case 4: return launch_b<4>(log2t, in, out, a, st);
        case 5: return launch_b<5>(log2t, in, out, a, st);
        case 6: return launch_b<6>(log2t, in, out, a, st);
        case 7: return launch_b<7>(log2t, in, out, a, st);
        case 8: return launch_b<8>(log2t, in, out, a, st);
        case 9: return launch_b<9>(log2t, in, out, a, st);
        case 10: return launch_b<10>(log2t, in, out, a, st);
        case 11: return launch_b<11>(log2t, in, out, a, st);
        case 12: return launch_b<12>(log2t, in, out, a, st);
        case 13: return launch_b<13>(log2t, in, out, a, st);
        default: return hipErrorInvalidValue;
    }
}

#if ESTL_PASS_DIR < 0
// the convolution's fused middle pass: last forward pass, spectrum multiply, first
// inverse pass (pass_tile.h, MID)
hipError_t launch_pass_mid(int log2b, int log2t, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st) {
    switch (log2b) {
        case 4: return launch_b<4, true>(log2t, in, out, a, st);
        case 5: return launch_b<5, true>(log2t, in, out, a, st);
        case 6: return launch_b<6, true>(log2t, in, out, a, st);
        case 7: return launch_b<7, true>(log2t, in, out, a, st);
        case 8: return launch_b<8, true>(log2t, in, out, a, st);
        case 9: return launch_b<9, true>(log2t, in, out, a, st);
        case 10: return launch_b<10, true>(log2t, in, out, a, st);
        case 11: return launch_b<11, true>(log2t, in, out, a, st);
        case 12: return launch_b<12, true>(log2t, in, out, a, st);
        case 13: return launch_b<13, true>(log2t, in, out, a, st);
        default: return hipErrorInvalidValue;
    }
}
#endif

}  // namespace estl
