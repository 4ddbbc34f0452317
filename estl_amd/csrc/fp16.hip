// fp16.hip -- FP16-storage / FP32-compute transforms with per-transform scaling.
//
// Data in HBM is IEEE binary16 (re, im) pairs -- half the bytes of complex64 -- and
// every butterfly runs in FP32 registers.  Scaling keeps every stored value in FP16's
// range with full precision:
//   * pack:   per transform t, m_t = max over its values of max(|re|, |im|); the
//             stored input is x / m_t (all components in [-1, 1]) and m_t is kept as
//             the transform's scale;
//   * passes: every pass of radix B multiplies what it stores by 1 / sqrt(B), so after
//             passes covering N_q points a component is at most sqrt(2 N_q) (<= 46341
//             for N <= 2^30, below FP16's 65504) while a random input keeps its RMS;
//   * the output's scale is the input's times sqrt(N) (times 1 / N when normalised):
//             true value = stored value * scale.
// The reference computes in FP32 only (FFT_OpenCL_Contiguous.h:47, cl_float2 buffers);
// this storage mode has no counterpart there.
#include "fft_block.h"
#include "pass_tile.h"

namespace estl {

// ------------------------------------------------------------------ scaling
__global__ void k_zero_u32(unsigned *__restrict__ p, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Max-abs per transform.  A workgroup of 256 threads takes a chunk of 2^c consecutive
// values, c = min(log2 n, 14): a piece of one transform, or whole transforms when
// n < 2^14.  Each thread reduces 2^(c-8) values (1 when c < 8); the threads of
// one transform then combine through xor shuffles inside the wave and, for chunks that
// span waves, through LDS; one atomicMax per transform and chunk (max is order-free,
// so the result does not depend on timing).
constexpr int kMaxAbsChunkLog2 = 14;

__global__ void __launch_bounds__(256) k_maxabs(const float2 *__restrict__ x, int log2n, long batch,
                                                unsigned *__restrict__ umax) {
    __shared__ float wmax[4];
    const int clog = log2n < kMaxAbsChunkLog2 ? log2n : kMaxAbsChunkLog2;  // chunk = min(n, 2^14)
    const int per_log = clog > 8 ? clog - 8 : 0;                            // values per thread
    const int tpt_log = log2n - per_log < 8 ? log2n - per_log : 8;          // threads per transform in this block
    const long base = (long)blockIdx.x << (per_log + 8);                    // first value of the block
    const long total = batch << log2n;
    const int t = threadIdx.x;
    float m = 0.0f;
    // thread t reads values base + t + 256 e (coalesced); a block of 2^(per_log + 8) > 256
    // values lies inside one transform, so all of them belong to first's transform
    const long first = base + t;
    if (first < total) {
        const float2 *p = x + first;
        for (int e = 0; e < (1 << per_log); ++e) {
            const float2 v = p[(long)e << 8];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
    }
    // lanes of one transform: groups of 2^tpt_log consecutive threads
    const int seg = tpt_log < 6 ? (1 << tpt_log) : 64;
    for (int off = 1; off < seg; off <<= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (tpt_log > 6) {  // the transform spans the block's waves
        if ((t & 63) == 0) wmax[t >> 6] = m;
        __syncthreads();
        if (t == 0) {
            float mm = wmax[0];
            for (int w = 1; w < (1 << (tpt_log - 6)); ++w) mm = fmaxf(mm, wmax[w]);
            if (first < total) atomicMax(umax + (first >> log2n), __float_as_uint(mm));
        }
        if (tpt_log == 7) {  // two transforms per block: the second half's leader
            if (t == 128 && first < total) atomicMax(umax + (first >> log2n), __float_as_uint(fmaxf(wmax[2], wmax[3])));
        }
    } else if (first < total && (t & (seg - 1)) == 0) {
        atomicMax(umax + (first >> log2n), __float_as_uint(m));
    }
}

// y = x / m_t in FP16; scale[t] = m_t (1 for an all-zero transform).  `scale` holds the
// atomicMax results (float bits) on entry and is rewritten in place with the same bits.
__global__ void k_pack(const float2 *__restrict__ x, f16x2 *__restrict__ y, int log2n, long batch,
                       float *__restrict__ scale) {
    const long count = batch << log2n;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const long t = i >> log2n;
        const float m = scale[t];
        const float d = m > 0.0f ? m : 1.0f;
        const float2 v = ld_stream(x + i);
        st_elem(y + i, make_float2(v.x / d, v.y / d));
        if ((i & ((1L << log2n) - 1)) == 0 && !(m > 0.0f)) scale[t] = 1.0f;
    }
}

__global__ void k_unpack(const f16x2 *__restrict__ y, const float *__restrict__ scale, float2 *__restrict__ x,
                         int log2n, long batch) {
    const long count = batch << log2n;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const float s = scale[i >> log2n];
        const float2 v = ld_elem(y + i);
        st_stream(x + i, make_float2(v.x * s, v.y * s));
    }
}

__global__ void k_scale_mul(const float *__restrict__ in, float *__restrict__ out, float f, long batch) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < batch) out[i] = in[i] * f;
}

static unsigned grid_for(long count, long per_block) {
    long b = (count + per_block - 1) / per_block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_pack_fp16(const float2 *x, void *y16, float *scale, int log2n, long batch, hipStream_t st) {
    f16x2 *y = static_cast<f16x2 *>(y16);
    unsigned *umax = reinterpret_cast<unsigned *>(scale);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(batch, 256)), dim3(256), 0, st, umax, batch);
    const int clog = log2n < kMaxAbsChunkLog2 ? log2n : kMaxAbsChunkLog2;
    const int per_log = clog > 8 ? clog - 8 : 0;
    const long blocks = ((batch << log2n) + (1L << (per_log + 8)) - 1) >> (per_log + 8);
    hipLaunchKernelGGL(k_maxabs, dim3((unsigned)blocks), dim3(256), 0, st, x, log2n, batch, umax);
    hipLaunchKernelGGL(k_pack, dim3(grid_for(batch << log2n, 256)), dim3(256), 0, st, x, y, log2n, batch, scale);
    return hipGetLastError();
}

hipError_t launch_unpack_fp16(const void *y16, const float *scale, float2 *x, int log2n, long batch, hipStream_t st) {
    const f16x2 *y = static_cast<const f16x2 *>(y16);
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(batch << log2n, 256)), dim3(256), 0, st, y, scale, x, log2n, batch);
    return hipGetLastError();
}

hipError_t launch_scale_mul(const float *in, float *out, float f, long batch, hipStream_t st) {
    hipLaunchKernelGGL(k_scale_mul, dim3(grid_for(batch, 256)), dim3(256), 0, st, in, out, f, batch);
    return hipGetLastError();
}

// ------------------------------------------------------------------ rows (N <= 2^13)
// One transform per TPT threads, TPW per workgroup; the workgroups walk the groups of
// rows with a stride of the grid, which the launcher sizes to one workgroup per group.
template <int LOG2N, int DIR, int TPW>
__global__ void __launch_bounds__(BlockGeom<LOG2N>::TPT * TPW)
k_fft_rows_h(const f16x2 *__restrict__ in, f16x2 *__restrict__ out, long batch, const float2 *__restrict__ tw,
             float scale) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    constexpr int SLICE = G::uses_lds ? Lds<kLdsPad>::size(N) : 0;
    constexpr int LDS = G::uses_lds ? TPW * SLICE : 1;
    __shared__ float2 lds[LDS];
    if constexpr (TPW == 1) {
        // one row per workgroup, the grid one workgroup per row (no clamp of the row
        // index or store guard: 69 / 77 VGPRs at 4096 / 8192 points instead of 75 / 83,
        // 1-5 % faster)
        const int tl = threadIdx.x;
        const long row = blockIdx.x;
        float2 v[E];
        const f16x2 *src = in + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = ld_elem(src + TPT * i);
        block_fft<LOG2N, DIR, kLdsPad, false, kRowsTw<LOG2N>>(v, lds, tl, tw, false);
        f16x2 *dst = out + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) st_elem(dst + TPT * i, cscale(v[i], scale));
        return;
    }
    const long groups = (batch + TPW - 1) / TPW;
    bool busy = false;
    for (long grp = blockIdx.x; grp < groups; grp += gridDim.x) {  // uniform across the workgroup
        const int t = opaque_lane(threadIdx.x);
        const int w = t / TPT, tl = t % TPT;
        const long want = grp * TPW + w;
        const long row = want < batch ? want : batch - 1;
        float2 v[E];
        const f16x2 *src = in + row * N + tl;
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = ld_elem(src + TPT * i);
        block_fft<LOG2N, DIR, kLdsPad, false, kRowsTw<LOG2N>>(v, lds + w * SLICE, tl, tw, busy);
        busy = true;
        if (want < batch) {
            f16x2 *dst = out + row * N + tl;
#pragma unroll
            for (int i = 0; i < E; ++i) st_elem(dst + TPT * i, cscale(v[i], scale));
        }
    }
}

template <int LOG2N, int DIR>
static hipError_t launch_rows_h_t(const f16x2 *in, f16x2 *out, long batch, const float2 *tw, float scale,
                                  hipStream_t st) {
    using G = BlockGeom<LOG2N>;
    constexpr int TPW = (G::TPT >= 256) ? 1 : 256 / G::TPT;
    auto kernel = k_fft_rows_h<LOG2N, DIR, TPW>;
    // one workgroup per group of rows (4-9 % faster than four resident rounds at 512 and
    // 1024 points, 1-3 % at 4096 and 8192)
    const long groups = (batch + TPW - 1) / TPW;
    if (groups > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, dim3((unsigned)groups), dim3(G::TPT * TPW), 0, st, in, out, batch, tw, scale);
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_rows_h_d(int log2n, const f16x2 *in, f16x2 *out, long batch, const float2 *tw, float scale,
                                  hipStream_t st) {
    switch (log2n) {
#define ROWS_H_CASE(L) \
    case L: return launch_rows_h_t<L, DIR>(in, out, batch, tw, scale, st);
        ROWS_H_CASE(1) ROWS_H_CASE(2) ROWS_H_CASE(3) ROWS_H_CASE(4) ROWS_H_CASE(5) ROWS_H_CASE(6) ROWS_H_CASE(7)
        ROWS_H_CASE(8) ROWS_H_CASE(9) ROWS_H_CASE(10) ROWS_H_CASE(11) ROWS_H_CASE(12) ROWS_H_CASE(13)
#undef ROWS_H_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows_fp16(int log2n, int dir, const void *in16, void *out16, long batch, const float2 *tw, float scale,
                            hipStream_t st) {
    const f16x2 *in = static_cast<const f16x2 *>(in16);
    f16x2 *out = static_cast<f16x2 *>(out16);
    return dir < 0 ? launch_rows_h_d<-1>(log2n, in, out, batch, tw, scale, st)
                   : launch_rows_h_d<+1>(log2n, in, out, batch, tw, scale, st);
}

// ------------------------------------------------------------------ global passes
template <int LOG2B, int DIR, int LOG2T, bool MID>
__global__ void __launch_bounds__(BlockGeom<LOG2B>::TPT << LOG2T)
k_fft_pass_h(const f16x2 *__restrict__ in, f16x2 *__restrict__ out, PassArgs a) {
    __shared__ float2 lds[PassTile<LOG2B, LOG2T>::lds_size];
    long tile = blockIdx.x;
    if (a.xcd_remap) {
        const long nb = gridDim.x;
        if ((nb & 7) == 0) tile = (tile & 7) * (nb >> 3) + (tile >> 3);
    }
    pass_tile<LOG2B, DIR, LOG2T, MID, f16x2>(in, out, a, tile, lds, false);
}

// FP16 passes: radices 2^4 .. 2^9 (every split the default planner makes), tiles of at
// most 2^14 values
constexpr int kMaxPassLog2Fp16 = 9;

template <int LOG2B, int DIR, int LOG2T, bool MID>
static hipError_t launch_pass_h_t(const f16x2 *in, f16x2 *out, const PassArgs &a, hipStream_t st) {
    if constexpr (LOG2B + LOG2T > kMaxPassTileLog2) {
        return hipErrorInvalidValue;
    } else {
        const int lcpr = a.log2n - LOG2B;
        if (lcpr < LOG2T) return hipErrorInvalidValue;
        const long blocks = a.batch << (lcpr - LOG2T);
        hipLaunchKernelGGL((k_fft_pass_h<LOG2B, DIR, LOG2T, MID>), dim3((unsigned)blocks), dim3(BlockGeom<LOG2B>::TPT << LOG2T),
                           0, st, in, out, a);
        return hipGetLastError();
    }
}

template <int LOG2B, int DIR, bool MID>
static hipError_t launch_pass_h_b(int log2t, const f16x2 *in, f16x2 *out, const PassArgs &a, hipStream_t st) {
    switch (log2t) {
        case 0: return launch_pass_h_t<LOG2B, DIR, 0, MID>(in, out, a, st);
        case 1: return launch_pass_h_t<LOG2B, DIR, 1, MID>(in, out, a, st);
        case 2: return launch_pass_h_t<LOG2B, DIR, 2, MID>(in, out, a, st);
        case 3: return launch_pass_h_t<LOG2B, DIR, 3, MID>(in, out, a, st);
        case 4: return launch_pass_h_t<LOG2B, DIR, 4, MID>(in, out, a, st);
        case 5: return launch_pass_h_t<LOG2B, DIR, 5, MID>(in, out, a, st);
        case 6: return launch_pass_h_t<LOG2B, DIR, 6, MID>(in, out, a, st);
        default: return hipErrorInvalidValue;
    }
}

template <int DIR, bool MID = false>
static hipError_t launch_pass_h_d(int log2b, int log2t, const f16x2 *in, f16x2 *out, const PassArgs &a,
                                  hipStream_t st) {
    switch (log2b) {
        case 4: return launch_pass_h_b<4, DIR, MID>(log2t, in, out, a, st);
        case 5: return launch_pass_h_b<5, DIR, MID>(log2t, in, out, a, st);
        case 6: return launch_pass_h_b<6, DIR, MID>(log2t, in, out, a, st);
        case 7: return launch_pass_h_b<7, DIR, MID>(log2t, in, out, a, st);
        case 8: return launch_pass_h_b<8, DIR, MID>(log2t, in, out, a, st);
        case 9: return launch_pass_h_b<9, DIR, MID>(log2t, in, out, a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pass_fp16(int log2b, int log2t, int dir, const void *in16, void *out16, const PassArgs &a,
                            hipStream_t st) {
    const f16x2 *in = static_cast<const f16x2 *>(in16);
    f16x2 *out = static_cast<f16x2 *>(out16);
    return dir < 0 ? launch_pass_h_d<-1>(log2b, log2t, in, out, a, st) : launch_pass_h_d<+1>(log2b, log2t, in, out, a, st);
}

// the FP16 convolution's middle pass: last forward pass, spectrum multiply, first
// inverse pass (pass_tile.h, MID)
hipError_t launch_pass_mid_fp16(int log2b, int log2t, const void *in16, void *out16, const PassArgs &a,
                                hipStream_t st) {
    return launch_pass_h_d<-1, true>(log2b, log2t, static_cast<const f16x2 *>(in16), static_cast<f16x2 *>(out16), a,
                                     st);
}

int max_pass_log2_fp16() { return kMaxPassLog2Fp16; }

}  // namespace estl
