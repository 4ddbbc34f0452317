// internal.h -- declarations shared between the kernels and the host-side plan code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace estl {

// Arguments of one global Stockham pass (k_fft_pass).
// A group of convolution requests of one length n run as one batch (the server's
// continuous batching): request q owns pair rows [pair_start[q], pair_start[q + 1]) of
// the batch.  Its real rows are read from x[q] and written to y[q] (real_rows[q] of
// them; FP32, or binary16 with FP16 storage), its spectrum is spec[q] (n bins); in the
// group's spectrum passes row q is the filter of request q (taps h[q], k[q] of them,
// rotation shift[q]).  FP16 storage: xscale[q] / yscale[q] are the member's input and
// output row scales and l1[q] its filter's (sum |h|, L) pair, filled on the device.
// Lives in device memory.
constexpr int kGroupMax = 16;
struct ConvGroup {
    int count;
    int pad_;
    long pair_start[kGroupMax + 1];
    long real_rows[kGroupMax];
    const void *x[kGroupMax];
    void *y[kGroupMax];
    const float2 *spec[kGroupMax];
    const float *h[kGroupMax];
    long k[kGroupMax];
    long shift[kGroupMax];
    const float *xscale[kGroupMax];
    float *yscale[kGroupMax];
    float2 l1[kGroupMax];
};

struct PassArgs {
    long batch;              // rows
    long A;                  // A_q: product of the radices of the earlier passes
    int log2n;               // row length n = 2^log2n (n <= 2^30)
    int log2S;               // S = n / (A_q B_q) = 2^log2S
    int lobits;              // two-level pre-twiddle split: m = (hi << lobits) | lo
    int xcd_remap;           // deal consecutive tiles to the same XCD
    float scale;             // applied to the stored values (1 = none)
    const float2 *tw_local;  // power table of the radix-B DFT (block_fft kTwPowers)
    const float2 *tw_lo;     // W_{A B}^{lo}, lo < 2^lobits
    const float2 *tw_hi;     // W_{A B}^{hi << lobits}
    // I/O modes of the convolution pipeline (0 = plain complex rows in and out):
    //   io_in  = 1: complex row r is the real pair (rin row 2r, rin row 2r + 1)
    //   spectrum / corr: the convolution middle pass (launch_pass_mid) multiplies its
    //   natural-order bins by spectrum[k] (its conjugate if corr)
    //   io_out = 2: write re / im of the scaled result to rout rows 2r, 2r + 1
    //   io_in  = 5 / io_out = 3 (request groups of single rows): complex row r is
    //               group member r's input x / output y (each one complex row)
    //   io_out = 4: the last pass of a request group's FP16 spectra: row r stored times
    //               1 / group->l1[r].x (member r's L), in place of the scale
    //   io_in  = 3: the row is a convolution filter built from its fk real taps at rin:
    //               folded modulo n and rotated by fshift (fold_tap); first pass of the
    //               filter's spectrum plan, direct load path (S >= T)
    //   io_in  = 4: the row is what a first pass of radix A makes of such a filter when
    //               its fk taps all lie within that pass's stride n / A, unrotated: its
    //               DFT of one nonzero value per column is that value in every output,
    //               so element e is tap e mod (n / A) (0 past fk); the second pass of the
    //               spectrum plan, which then skips its first pass (either load path)
    int io_in, io_out, corr;
    long real_rows;          // real rows behind the pairs (the last pair may be single)
    const float *rin;
    float *rout;
    const void *rin16;       // the same, binary16 reals (FP16-storage passes)
    void *rout16;
    int planar_in;           // `in` / `out` use the tile-planarized layout (FP32 only)
    int planar_out;
    const float2 *spectrum;
    long fk, fshift;         // io_in = 3, 4: taps and rotation of the filter
    // convolution passes of a request group: rows index the group's batch, and the
    // real-pair I/O, the filter loads and the middle pass's spectrum are the member's
    const ConvGroup *group;
};

hipError_t launch_rows(int log2n, int dir, const float2 *in, float2 *out, long batch, const float2 *tw, float scale,
                       hipStream_t st);
// workgroups of a kernel co-resident on the current device (cached per kernel and
// device, thread-safe; kernels.hip)
long co_resident_blocks(const void *kernel, int block_threads);
// rows kernel on the tile-planarized layout (kernels.hip)
hipError_t launch_rows_planar(int log2n, int dir, const float2 *in, float2 *out, long batch, const float2 *tw,
                              float scale, hipStream_t st);
hipError_t launch_pass_fwd(int log2b, int log2t, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st);
hipError_t launch_pass_bwd(int log2b, int log2t, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st);
// FP16 storage (fp16.hip): buffers of (re, im) binary16 pairs, passed untyped
hipError_t launch_pack_fp16(const float2 *x, void *y16, float *scale, int log2n, long batch, hipStream_t st);
hipError_t launch_unpack_fp16(const void *y16, const float *scale, float2 *x, int log2n, long batch, hipStream_t st);
hipError_t launch_scale_mul(const float *in, float *out, float f, long batch, hipStream_t st);
hipError_t launch_rows_fp16(int log2n, int dir, const void *in16, void *out16, long batch, const float2 *tw, float scale,
                            hipStream_t st);
hipError_t launch_pass_fp16(int log2b, int log2t, int dir, const void *in16, void *out16, const PassArgs &a,
                            hipStream_t st);
hipError_t launch_pass_mid_fp16(int log2b, int log2t, const void *in16, void *out16, const PassArgs &a,
                                hipStream_t st);
int max_pass_log2_fp16();
// forward last pass + spectrum multiply + inverse first pass of a convolution
hipError_t launch_pass_mid(int log2b, int log2t, const float2 *in, float2 *out, const PassArgs &a, hipStream_t st);
inline hipError_t launch_pass(int log2b, int log2t, int dir, const float2 *in, float2 *out, const PassArgs &a,
                              hipStream_t st) {
    return dir < 0 ? launch_pass_fwd(log2b, log2t, in, out, a, st) : launch_pass_bwd(log2b, log2t, in, out, a, st);
}
hipError_t launch_twiddle(float2 *dst, long count, int mode, long A, long mult, long nq, hipStream_t st);
hipError_t launch_fill_uniform(float *dst, long count, uint64_t seed, uint64_t offset, hipStream_t st);
// copy n16 blocks of 16 bytes (16-byte aligned), nontemporal or plain loads and stores
hipError_t launch_stream_copy(const void *src, void *dst, long n16, bool nontemporal, hipStream_t st);

long rows_twiddle_count(int log2n);
long rows_power_count(int log2n);  // entries of BlockGeom<log2n>'s power table
void rows_pass_params(int log2n, int q, long *A, int *B);
int rows_pass_count(int log2n);

constexpr int kMaxRowsLog2 = 13;   // largest single-launch row length
constexpr int kSpecBlockTwLog2 = 12;  // filter spectra of at least 2^12 points read the block table
constexpr int kMaxPassLog2 = 13;   // largest radix of a global pass
constexpr int kMaxPassTileLog2 = 14;  // a pass tile holds at most 2^14 values (~139 KiB of LDS)
constexpr int kMaxPassCols = 6;    // at most 2^6 columns per tile
constexpr int kOlsLog2 = 13;       // overlap-save convolution windows of 2^13 values
constexpr int kOlsWideLog2 = 14;   // and of 2^14 values (rows of n >= 2^15)
#ifndef ESTL_OLS_MAX_TAPS
#define ESTL_OLS_MAX_TAPS 4097
#endif
constexpr long kOlsMaxTaps = ESTL_OLS_MAX_TAPS;  // filters up to this length fit 2^13 windows
constexpr long kOlsWideMaxTaps = 8193;           // and up to this length 2^14 windows
// Overlap-save windows start on 128-byte boundaries (32 FP32 values, 64 binary16 ones):
// the first valid position of a window moves from pre (K - 1, or 0 for a correlation) up
// to the next multiple of the alignment, and the outputs per window are the multiple of
// it that still fits.
inline long ols_align(bool fp16) { return fp16 ? 64 : 32; }
inline long ols_first_valid(long k, bool corr, bool fp16 = false) {
    const long pre = corr ? 0 : k - 1, al = ols_align(fp16);
    return (pre + al - 1) / al * al;
}
inline long ols_outputs(int log2L, long k, bool corr, bool fp16 = false) {
    const long pre = corr ? 0 : k - 1, al = ols_align(fp16);
    return ((1L << log2L) - k + 1 - (ols_first_valid(k, corr, fp16) - pre)) / al * al;
}

}  // namespace estl
