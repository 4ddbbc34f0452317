// fft_block.h -- workgroup-cooperative Stockham FFT held in registers, exchanged
// through LDS.
//
// A sequence of length N = 2^LOG2N is spread over TPT = N / E threads, E = 16 values
// per thread: thread tl holds element tl + TPT*i in v[i].  The plan is the reference's
// global radix list (FFT_OpenCL_Contiguous.h:55-60: radix-16 passes, the 2^(log2N mod 4)
// remainder last) and every pass follows its Stockham recurrence
// (FFT_OpenCL_Contiguous.h:167-180, docs/ImplementationStrategy.pdf eq. 7/12):
//
//   d_q[(N/N_q)(g + A_q h) + z] = sum_s d_{q-1}[(N/N_q)(B_q g + s) + z] W_{N_q}^{g s} W_{B_q}^{h s}
//
// but where the reference runs one kernel launch (one HBM round trip) per pass, here a
// pass is a register DFT and the hand-over between passes is an LDS exchange.  After
// every pass thread tl holds d_q[tl + TPT*i] in v[i] again, so the first pass needs no
// exchange (its inputs are exactly what a coalesced load gives) and the last pass
// leaves the result in the order a coalesced store wants.
#pragma once

#include "fft_device.h"
#include "internal.h"

namespace estl {

template <int LOG2N>
struct BlockGeom {
    static constexpr int N = 1 << LOG2N;
    static constexpr int E = (LOG2N >= 4) ? 16 : N;  // complex values per thread
    static constexpr int TPT = N / E;                // threads per sequence
    static constexpr int NFULL = LOG2N / 4;          // radix-16 passes
    static constexpr int REM = LOG2N % 4;
    static constexpr int P = NFULL + (REM ? 1 : 0);  // passes
    static constexpr int radix(int q) { return q < NFULL ? 16 : (1 << REM); }
    static constexpr long A(int q) {
        long a = 1;
        for (int p = 0; p < q; ++p) a *= radix(p);
        return a;
    }
    // offset of pass q's twiddles in the table; pass q stores W_{N_q}^{g s} at
    // [(s - 1) * A_q + g] for s in [1, B_q), g in [0, A_q) (pass 0 needs none)
    static constexpr long twoff(int q) {
        long o = 0;
        for (int p = 1; p < q; ++p) o += (radix(p) - 1) * A(p);
        return o;
    }
    static constexpr long twsize = twoff(P);
    // the power table (TwPowers): W_{N_q}^g for g in [0, A_q) per pass q >= 1, at
    // powoff(q)
    static constexpr long powoff(int q) {
        long o = 0;
        for (int p = 1; p < q; ++p) o += A(p);
        return o;
    }
    static constexpr long powsize = powoff(P);
    static constexpr bool uses_lds = TPT > 1;
};

// Twiddle sources of block_fft's passes q >= 1 (each butterfly of pass q, group g, needs
// W_{N_q}^{g s} for s in [1, B_q)):
//   kTwBlock:  every W_{N_q}^{g s} read from the block table (B_q - 1 reads per
//              butterfly; the table is N values, 128 KiB at N = 2^14);
//   kTwPowers: one read, w = W_{N_q}^g from the power table (sum of the A_q, 34 KiB at
//              2^14, mostly cache-resident); w^(2^k) by squaring, and w^s as the product
//              of the w^(2^k) of the set bits of s -- at most six products deep from the
//              correctly rounded w (a few ulp), with only the log2 B powers live.  Trades
//              the long-latency reads of a table too big for L1 for VALU work.
enum { kTwBlock = 0, kTwPowers = 1 };
// Twiddle source of the rows kernels (k_fft_rows, _planar, _h): powers at every
// row length (a rows plan's table is the power table, estl_c.cpp).  With powers the
// rows kernel of 4096 and 8192 points needs 68-75 VGPRs; the block table's twiddles
// kept in registers across rows had needed 167-206 (three waves per SIMD at 4096 points,
// two at 8192) and measured 6-9 % slower, and read again for each row from L2 they cost
// more.  The price is a few ulp per twiddle: rel-L2 against FP64 1.2e-7 -> 3.0e-7 at
// 4096 points (the bound is 1.2e-5).
template <int LOG2N>
constexpr int kRowsTw = kTwPowers;
// Twiddle source of the filter spectrum (k_filter_spectrum): the block table from 4096
// points.  A spectrum is computed once per request or registration, so its speed hardly
// matters, and the correctly rounded twiddles keep, e.g., a delayed unit impulse's
// spectrum -- and the convolution by it -- within 2e-6 of the shifted input.
template <int LOG2N>
constexpr int kSpecTw = LOG2N < kSpecBlockTwLog2 ? kTwPowers : kTwBlock;
template <int B, int DIR>
__device__ __forceinline__ void apply_twiddle_powers(float2 (&u)[B], float2 w) {
    constexpr int LB = __builtin_ctz(B);
    float2 p[LB > 0 ? LB : 1];
    p[0] = w;
    StaticFor<1, LB>::run([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        p[k] = cmul(p[k - 1], p[k - 1]);
    });
    StaticFor<1, B>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int lo = __builtin_ctz(s);
        float2 t = p[lo];
        StaticFor<lo + 1, LB>::run([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            if constexpr ((s >> b) & 1) t = cmul(t, p[b]);
        });
        u[s] = ctw<DIR>(u[s], t);
    });
}

// LDS layouts of a sequence's exchange buffer.
//  kLdsPad: one float2 of padding after every 16 values; addresses stay base + constant,
//           so the exchange reads use immediate offsets (fewest VGPRs);
//  kLdsXor: no padding, value i sits in row i/16 at column (i mod 16) XOR (row mod 16).
//           Every exchange of these passes touches, per 16 lanes, either 16 consecutive
//           values of one row or one value from each of 16 rows at a fixed column; the
//           XOR spreads both over 16 distinct columns (all 32 four-byte banks for 8-byte
//           values) and a 4096-value sequence needs exactly 32 KiB.
enum { kLdsPad = 0, kLdsXor = 1 };
template <int LAYOUT>
struct Lds;
template <>
struct Lds<kLdsPad> {
    static constexpr int size(int n) { return n + (n >> 4); }
    static __device__ __forceinline__ int idx(int i) { return i + (i >> 4); }
};
template <>
struct Lds<kLdsXor> {
    static constexpr int size(int n) { return n; }
    static __device__ __forceinline__ int idx(int i) { return i ^ ((i >> 4) & 15); }
};

// On entry v[i] = x[tl + TPT*i]; on exit v[i] = X[tl + TPT*i] (X = DFT of x, sign DIR).
// `lds` is this sequence's LDS slice (Lds<LAYOUT>::size(N) float2); `tw` the size-N block
// table (TWSRC = kTwBlock) or the power table (kTwPowers).
// All threads of the workgroup must call it together (it holds workgroup barriers).
// If `lds_busy` is set, the slice is still being read on entry and a barrier is placed
// before the first LDS write.  HOIST_TW lets the compiler keep the twiddles of every pass
// in registers across calls made from a loop over one size; callers that dispatch many
// sizes in one loop pass false so each size's twiddles are reloaded (from L1/L2)
// instead of all staying live.  No kernel sets it at present: the rows kernels, its
// last user, run faster at one row per workgroup with twiddles by powers (k_fft_rows).
template <int LOG2N, int DIR, int LAYOUT = kLdsPad, bool HOIST_TW = false, int TWSRC = kTwBlock>
__device__ __forceinline__ void block_fft(float2 (&v)[BlockGeom<LOG2N>::E], float2 *lds, int tl,
                                          const float2 *__restrict__ tw, bool lds_busy = false) {
    using G = BlockGeom<LOG2N>;
    constexpr int N = G::N, E = G::E, TPT = G::TPT;
    StaticFor<0, G::P>::run([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int B = G::radix(q);
        constexpr int NB = E / B;  // butterflies per thread in this pass
        constexpr int Aq = int(G::A(q));
        constexpr int S = N / (Aq * B);
        if constexpr (q == 0) {
            StaticFor<0, NB>::run([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                float2 u[B];
                StaticFor<0, B>::run([&](auto sc) { u[decltype(sc)::value] = v[b + NB * decltype(sc)::value]; });
                Dft<B, DIR>::run(u);
                StaticFor<0, B>::run([&](auto hc) { v[b + NB * decltype(hc)::value] = u[decltype(hc)::value]; });
            });
        } else {
            if (q > 1 || lds_busy) __syncthreads();
#pragma unroll
            for (int i = 0; i < E; ++i) lds[Lds<LAYOUT>::idx(tl + TPT * i)] = v[i];
            __syncthreads();
            StaticFor<0, NB>::run([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                const int j = tl + TPT * b;
                const int g = j / S, z = j % S;   // S is a power of two
                const int base = S * B * g + z;
                float2 u[B];
#pragma unroll
                for (int s = 0; s < B; ++s) u[s] = lds[Lds<LAYOUT>::idx(base + S * s)];
                if constexpr (TWSRC == kTwPowers) {
                    apply_twiddle_powers<B, DIR>(u, (HOIST_TW ? tw : opaque(tw))[G::powoff(q) + g]);
                } else {
                    const float2 *twq = (HOIST_TW ? tw : opaque(tw)) + G::twoff(q) + g;
#pragma unroll
                    for (int s = 1; s < B; ++s) u[s] = ctw<DIR>(u[s], twq[(s - 1) * Aq]);
                }
                Dft<B, DIR>::run(u);
#pragma unroll
                for (int h = 0; h < B; ++h) v[b + NB * h] = u[h];
            });
        }
    });
}

}  // namespace estl
