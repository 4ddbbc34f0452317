// pass_tile.h -- one tile of a global Stockham pass, shared by k_fft_pass (one tile
// per workgroup) and the uneven-batch work-queue kernel (many tiles per workgroup).
//
// Pass q over `batch` rows of length n = 2^log2n with radix B = 2^LOG2B, A = A_q,
// S = n / (A B) = 2^log2S follows FFT_OpenCL_Contiguous.h:167-180: column
// j in [0, n/B) of a row has g = j / S, z = j mod S and
//   load   x_s = in[S (B g + s) + z] * W_{A B}^{g s}      s in [0, B)
//   DFT_B  of x (block_fft: registers + LDS exchanges)
//   store  out[j + h n / B] = X_h                          h in [0, B)
// Where the reference launches one kernel per radix-16 pass, a pass here is a radix of
// up to 2^13 computed inside one workgroup.  A tile is T = 2^LOG2T adjacent columns
// (thread c + T * tl serves column c), so every load and store moves T consecutive
// complex values: when S >= T the loads read T contiguous values per s; otherwise the
// tile's input is one contiguous block of T B values, staged through LDS.  Offsets inside
// a row are 32-bit (n <= 2^30); input and output are streamed.
#pragma once

#include "fft_block.h"
#include "internal.h"

namespace estl {

// exact pre-twiddle table reads every kTwRefresh-th element of a thread (see pass_tile)
#ifndef ESTL_TW_REFRESH
#define ESTL_TW_REFRESH 4
#endif
constexpr int kTwRefresh = ESTL_TW_REFRESH;

template <int LOG2B, int LOG2T>
struct PassTile {
    static constexpr int B = 1 << LOG2B, T = 1 << LOG2T;
    // Adjacent lanes serve adjacent columns, so column slices are strided by an odd
    // number of float2: 16 lanes at the same offset in 16 columns then land on 16
    // distinct 8-byte bank pairs (all 32 banks).
    static constexpr int ldsp = Lds<kLdsPad>::size(B) | 1;
    static constexpr int lds_size = T * ldsp;
    static constexpr int threads = BlockGeom<LOG2B>::TPT << LOG2T;
};

// The member of a request group that owns pair row `row` (uniform across the tile).
__device__ __forceinline__ int group_member(const ConvGroup *g, long row) {
    int q = 0;
    while (q + 1 < g->count && g->pair_start[q + 1] <= row) ++q;
    return q;
}

// Geometry of tile `tile` for the calling thread: the row, the tile's first column j0,
// the thread's column j = j0 + c, its g = j / S and z = j mod S, and tl.
template <int LOG2B, int LOG2T>
struct TilePos {
    long row;
    int j0, j, g, z, c, tl, t;
    __device__ __forceinline__ TilePos(const PassArgs &a, long tile) {
        t = opaque_lane(threadIdx.x);
        c = t & ((1 << LOG2T) - 1);
        tl = t >> LOG2T;
        const int lcpr = a.log2n - LOG2B;  // columns per row = 2^lcpr
        row = tile >> (lcpr - LOG2T);
        j0 = int(tile - (row << (lcpr - LOG2T))) << LOG2T;
        j = j0 + c;
        g = j >> a.log2S;
        z = j & ((1 << a.log2S) - 1);
    }
};

// Phase 1 of a tile: its HBM loads, E values per thread.  Direct path (S >= T) and the
// real-pair path: the thread's own column, element s = tl + TPT i in v[i].  Staged path
// (S < T): the thread's share of the tile's input, a contiguous block of T B values
// starting at B j0 (value t + NT k in v[k]); pass_tile_finish routes it through LDS.
template <int LOG2B, int LOG2T, typename ST = float2>
__device__ __forceinline__ void pass_tile_load(const ST *__restrict__ in, const PassArgs &a, long tile,
                                               float2 (&v)[BlockGeom<LOG2B>::E]) {
    constexpr bool F32 = __is_same(ST, float2);
    using G = BlockGeom<LOG2B>;
    constexpr int E = G::E, TPT = G::TPT, NT = TPT << LOG2T;
    const TilePos<LOG2B, LOG2T> P(a, tile);
    const int ls = a.log2S;
    if (a.io_in == 1) {
        // real pair: rows 2 row and 2 row + 1 of rin (the second absent past the end),
        // FP32 or, with FP16 storage, binary16 reals; only the first pass reads pairs
        // and its plans keep S >= T.  In a request group: the member's pair and rows.
        long prow = P.row, rrows = a.real_rows;
        const void *xb = F32 ? static_cast<const void *>(a.rin) : a.rin16;
        if (a.group) {
            const int q = group_member(a.group, P.row);
            prow = P.row - a.group->pair_start[q];
            rrows = a.group->real_rows[q];
            xb = a.group->x[q];
        }
        const bool two = 2 * prow + 1 < rrows;
        const int base = (P.g << (ls + LOG2B)) + P.z;
        if constexpr (F32) {
            const float *x1 = static_cast<const float *>(xb) + ((2 * prow) << a.log2n);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int o = base + ((P.tl + TPT * i) << ls);
                v[i] = make_float2(ld_stream_f(x1 + o), two ? ld_stream_f(x1 + (1L << a.log2n) + o) : 0.0f);
            }
        } else {
            const _Float16 *x1 = static_cast<const _Float16 *>(xb) + ((2 * prow) << a.log2n);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int o = base + ((P.tl + TPT * i) << ls);
                v[i] = make_float2((float)x1[o], two ? (float)x1[(1L << a.log2n) + o] : 0.0f);
            }
        }
    } else if (F32 && a.io_in == 3) {
        // the filter, built from its taps as it is loaded; its plans keep S >= T.  In a
        // request group, row q is member q's filter
        const long n = 1L << a.log2n;
        const int base = (P.g << (ls + LOG2B)) + P.z;
        const float *h = a.rin;
        long fk = a.fk, fsh = a.fshift;
        if (a.group) {
            h = a.group->h[P.row];
            fk = a.group->k[P.row];
            fsh = a.group->shift[P.row];
        }
#pragma unroll
        for (int i = 0; i < E; ++i)
            v[i] = make_float2(fold_tap(h, fk, n, fsh, base + ((P.tl + TPT * i) << ls)), 0.0f);
    } else if (F32 && a.io_in == 4) {
        // the first spectrum pass's output, rebuilt from the taps (the host checks that
        // every tap lies within the first pass's stride n / A): element e is tap
        // e mod (n / A), read at the element indices of the direct or the staged layout
        const float *h = a.rin;
        long fk = a.fk;
        if (a.group) {
            h = a.group->h[P.row];
            fk = a.group->k[P.row];
        }
        const int mask = int(((1L << a.log2n) / a.A) - 1);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int e = ls >= LOG2T ? (P.g << (ls + LOG2B)) + P.z + ((P.tl + TPT * i) << ls)
                                      : (P.j0 << LOG2B) + P.t + NT * i;
            const int m = e & mask;
            v[i] = make_float2(m < fk ? h[m] : 0.0f, 0.0f);
        }
    } else if (F32 && a.planar_in) {
        // the same values from the tile-planarized layout, by element index
        const long e0 = (P.row << a.log2n) +
                        (ls >= LOG2T ? long((P.g << (ls + LOG2B)) + P.z) : long(P.j0 << LOG2B) + P.t);
#pragma unroll
        for (int i = 0; i < E; ++i)
            v[i] = ld_planar(reinterpret_cast<const float2 *>(in),
                             e0 + (ls >= LOG2T ? long((P.tl + TPT * i) << ls) : long(NT * i)));
    } else {
        // a plain complex row (io_in 5: a request group member's, one row each)
        const ST *rowp = in + (P.row << a.log2n);
        if constexpr (F32)
            if (a.io_in == 5) rowp = static_cast<const ST *>(a.group->x[P.row]);
        if (ls >= LOG2T) {
            const ST *p = rowp + ((P.g << (ls + LOG2B)) + P.z);
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = ld_elem(p + ((P.tl + TPT * i) << ls));
        } else {
            const ST *p = rowp + (P.j0 << LOG2B);
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = ld_elem(p + P.t + NT * k);
        }
    }
}

// Phase 2 of a tile: (staged path: the LDS exchange that puts each column's values in
// its threads,) pre-twiddle, radix-B DFT, store.  All threads of the workgroup
// (PassTile::threads of them) call it together.  `lds` holds PassTile::lds_size
// float2; if `lds_busy`, a previous user of it may still be reading and a barrier
// comes first.
//
// MID (forward only): the last forward pass of a convolution fused with the first
// inverse pass.  The last forward pass (S = 1) of column j stores bin j + h n/B; the
// first pass of an inverse whose radices run in the reverse order (A = 1, S = n/B, no
// pre-twiddle) loads exactly x_s = in[j + s n/B] for its column j.  So the tile
// multiplies its B bins by the spectrum and runs the inverse radix-B DFT in registers,
// saving one HBM round trip and one launch.
//
// ST is the storage type of `in` and `out`: float2, or f16x2 (FP16 storage, FP32
// arithmetic; its real-pair I/O modes read and write binary16 reals, rin16 / rout16).
template <int LOG2B, int DIR, int LOG2T, bool MID = false, typename ST = float2>
__device__ __forceinline__ void pass_tile_finish(ST *__restrict__ out, const PassArgs &a, long tile,
                                                 float2 (&v)[BlockGeom<LOG2B>::E], float2 *lds, bool lds_busy) {
    constexpr bool F32 = __is_same(ST, float2);
    using G = BlockGeom<LOG2B>;
    constexpr int E = G::E, TPT = G::TPT, T = 1 << LOG2T, NT = T * TPT;
    constexpr int LDSP = PassTile<LOG2B, LOG2T>::ldsp;
    const TilePos<LOG2B, LOG2T> P(a, tile);
    const int ls = a.log2S;
    const int lcpr = a.log2n - LOG2B;
    const int c = P.c, tl = P.tl, j = P.j;
    bool busy = lds_busy;
    if ((a.io_in == 0 || a.io_in == 4 || a.io_in == 5) && ls < LOG2T) {
        // value o of the contiguous block belongs to column (o / (S B)) S + o mod S,
        // element (o mod S B) / S
        const int sbmask = (1 << (ls + LOG2B)) - 1, smask = (1 << ls) - 1;
        if (busy) __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int o = P.t + NT * k;
            const int gl = o >> (ls + LOG2B), rem = o & sbmask;
            const int s = rem >> ls, zz = rem & smask;
            lds[((gl << ls) + zz) * LDSP + Lds<kLdsPad>::idx(s)] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E; ++i) v[i] = lds[c * LDSP + Lds<kLdsPad>::idx(tl + TPT * i)];
        busy = true;
    }
    if (a.A > 1) {
        // W_{A B}^{g s} for s = tl + TPT i.  A table value is W^{lo} W^{hi << lobits},
        // m = g s < A B (two exactly rounded entries, one complex product).  Exact
        // values are taken at i = 0, R, 2R, ...; in between, w_{kR + r} = w_{kR} step^r
        // with step = W^{g TPT} -- at most two extra roundings per twiddle, and
        // 2 (E / R + 1) table reads per thread instead of 2 E.
        const unsigned lomask = (1u << a.lobits) - 1;
        const unsigned gu = unsigned(P.g);
        auto table = [&](unsigned m) { return cmul(a.tw_lo[m & lomask], a.tw_hi[m >> a.lobits]); };
        constexpr int R = (kTwRefresh < E) ? kTwRefresh : E;
        float2 sp[R];
        sp[0] = make_float2(1.0f, 0.0f);
        if constexpr (R > 1) {
            sp[1] = table(gu * unsigned(TPT));
#pragma unroll
            for (int r = 2; r < R; ++r) sp[r] = cmul(sp[r >> 1], sp[r - (r >> 1)]);
        }
        float2 base = sp[0];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            if (i % R == 0) base = table(gu * unsigned(tl + TPT * i));
            const float2 w = (i % R == 0) ? base : cmul(base, sp[i % R]);
            v[i] = ctw<DIR>(v[i], w);
        }
    }
    block_fft<LOG2B, DIR, kLdsPad, false, kTwPowers>(v, lds + c * LDSP, tl, a.tw_local, busy);
    if (a.io_out == 2) {
        // re / im to real rows 2 row and 2 row + 1 of rout (FP32 or binary16); in a
        // request group, the member's pair and rows
        long prow = P.row, rrows = a.real_rows;
        void *yb = F32 ? static_cast<void *>(a.rout) : a.rout16;
        if (a.group) {
            const int q = group_member(a.group, P.row);
            prow = P.row - a.group->pair_start[q];
            rrows = a.group->real_rows[q];
            yb = a.group->y[q];
        }
        const bool two = 2 * prow + 1 < rrows;
        if constexpr (F32) {
            float *y1 = static_cast<float *>(yb) + ((2 * prow) << a.log2n) + j;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int o = (tl + TPT * i) << lcpr;
                st_stream_f(y1 + o, v[i].x * a.scale);
                if (two) st_stream_f(y1 + (1L << a.log2n) + o, v[i].y * a.scale);
            }
        } else {
            _Float16 *y1 = static_cast<_Float16 *>(yb) + ((2 * prow) << a.log2n) + j;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int o = (tl + TPT * i) << lcpr;
                y1[o] = (_Float16)(v[i].x * a.scale);
                if (two) y1[(1L << a.log2n) + o] = (_Float16)(v[i].y * a.scale);
            }
        }
        return;
    }
    if constexpr (MID) {
        // the last forward pass yields natural order: output h of column j is bin j + h n/B
        const float2 *spec = a.group ? a.group->spec[group_member(a.group, P.row)] : a.spectrum;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const float2 hk = spec[j + ((tl + TPT * i) << lcpr)];
            v[i] = a.corr ? cmulc(v[i], hk) : cmul(v[i], hk);
        }
        block_fft<LOG2B, -DIR, kLdsPad, false, kTwPowers>(v, lds + c * LDSP, tl, a.tw_local, true);
    }
    if (F32 && a.planar_out) {
        const long e0 = (P.row << a.log2n) + j;
#pragma unroll
        for (int i = 0; i < E; ++i)
            st_planar(reinterpret_cast<float2 *>(out), e0 + ((long)(tl + TPT * i) << lcpr), cscale(v[i], a.scale));
        return;
    }
    ST *dst = out + (P.row << a.log2n) + j;
    float sc = a.scale;
    if constexpr (F32) {
        if (a.io_out == 3) dst = static_cast<ST *>(a.group->y[P.row]) + j;
        if (a.io_out == 4) sc = 1.0f / a.group->l1[P.row].x;
    }
#pragma unroll
    for (int i = 0; i < E; ++i) st_elem(dst + ((tl + TPT * i) << lcpr), cscale(v[i], sc));
}

// Computes tile `tile` of the pass described by `a` (load, then finish).
template <int LOG2B, int DIR, int LOG2T, bool MID = false, typename ST = float2>
__device__ __forceinline__ void pass_tile(const ST *__restrict__ in, ST *__restrict__ out, const PassArgs &a,
                                          long tile, float2 *lds, bool lds_busy) {
    float2 v[BlockGeom<LOG2B>::E];
    pass_tile_load<LOG2B, LOG2T, ST>(in, a, tile, v);
    pass_tile_finish<LOG2B, DIR, LOG2T, MID, ST>(out, a, tile, v, lds, lds_busy);
}

}  // namespace estl
