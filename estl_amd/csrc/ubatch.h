// ubatch.h -- descriptors of the uneven-size batch (work-queue kernel k_ubatch).
#pragma once

#include <hip/hip_runtime.h>

namespace estl {

// work-queue workgroup size 2^kUThreadsLog2 (see workqueue.hip)
#ifndef ESTL_UBATCH_TLOG
#define ESTL_UBATCH_TLOG 9
#endif
constexpr int kUThreadsLog2 = ESTL_UBATCH_TLOG;
constexpr int kUMaxLog2 = 20;  // largest transform of an uneven batch
// up to here a work item is a group of rows (a large item must fill a workgroup's tile:
// 2^(kUThreadsLog2 + 4) values)
constexpr int kUMaxSmallLog2 = kUThreadsLog2 + 3 > 12 ? kUThreadsLog2 + 3 : 12;

// One work item: `count` transforms of size 2^log2n, the `first`-th.. of the
// size-sorted offset list (count = 1 for large sizes).
struct UItem {
    int log2n;
    int count;
    int first;
    int pad;
};

// Device tables of every size class.
struct UTables {
    // power tables (block_fft kTwPowers) of BlockGeom<log2>, log2 0..kUMaxSmallLog2: the
    // small transforms and the radices of the large ones' passes
    const float2 *pow[kUMaxSmallLog2 + 1];
    const float2 *lo[kUMaxLog2 + 1];          // pre-twiddle tables of the second pass
    const float2 *hi[kUMaxLog2 + 1];
    int lobits[kUMaxLog2 + 1];
};

struct UArgs {
    const UItem *items;
    int n_items;
    const long *offs;     // element offset of each transform, size-sorted
    const float2 *in;
    float2 *out;
    float2 *scratch;      // pass-1 output of large transforms (same offsets)
    int *counters;        // [0] ticket, [1] workgroups done (both zero between launches)
    float scale;
    UTables tw;
};

hipError_t launch_ubatch(int dir, const UArgs &a, long blocks, hipStream_t st);
long ubatch_resident_blocks();
int ubatch_small_rows_per_item(int log2n);

}  // namespace estl
