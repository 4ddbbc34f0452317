// ubatch_c.cpp -- C ABI of the uneven-size batch (k_ubatch work-queue kernel).
//
// create: validate the (size, offset) descriptors, sort the transforms by size
//         (largest first), cut them into work items (groups of same-size rows for
//         n <= 2^12, one transform per item above), upload the item list, the sorted
//         offsets and the twiddle-table pointers, allocate the scratch used by the
//         two-pass transforms and the two ticket counters (zeroed once; the kernel
//         leaves them zero);
// execute: one launch of k_ubatch on the caller's stream;
// execute_graph: the same launch captured once in a HIP graph per (in, out, stream)
//         and replayed (no memset node is recorded: the kernel resets its counters).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/estl_hip.h"
#include "plan.h"
#include "ubatch.h"

namespace estl {

struct UBatch {
    int device = 0, dir = -1, count = 0;
    unsigned flags = 0;
    std::vector<long> sizes, offsets;
    std::vector<UItem> items;
    UItem *d_items = nullptr;
    long *d_offs = nullptr;
    int *d_counters = nullptr;
    float2 *scratch = nullptr;
    size_t scratch_elems = 0;
    long blocks = 0;
    float scale = 1.0f;  // only used when every transform has the same size
    UArgs args{};
    // captured graph for execute_graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const void *g_in = nullptr;
    void *g_out = nullptr;
    hipStream_t cap_stream = nullptr;  // private stream the graph is captured on
    long total_elems = 0, large_elems = 0, extent = 0;
};

static int ilog2l(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

static void ubatch_free(UBatch *u) {
    if (!u) return;
    if (u->gexec) (void)hipGraphExecDestroy(u->gexec);
    if (u->graph) (void)hipGraphDestroy(u->graph);
    if (u->cap_stream) (void)hipStreamDestroy(u->cap_stream);
    if (u->d_items) (void)hipFree(u->d_items);
    if (u->d_offs) (void)hipFree(u->d_offs);
    if (u->d_counters) (void)hipFree(u->d_counters);
    if (u->scratch) (void)hipFree(u->scratch);
    delete u;
}

#define HIP_TRY_U(expr)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            ubatch_free(u);                                                                           \
            return set_error(ESTL_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));            \
        }                                                                                             \
    } while (0)

static int ubatch_create(UBatch **out, int count, const long *sizes, const long *offsets, int direction,
                         unsigned flags) {
    *out = nullptr;
    if (count < 1 || !sizes) return set_error(ESTL_ERR_INVALID, "empty uneven batch");
    if (direction != ESTL_FORWARD && direction != ESTL_BACKWARD)
        return set_error(ESTL_ERR_INVALID, "direction=%d must be -1 or +1", direction);
    if (flags & ~(unsigned)ESTL_FLAG_NORMALIZE)
        return set_error(ESTL_ERR_UNSUPPORTED, "flags 0x%x not supported for uneven batches", flags);
    UBatch *u = new UBatch();
    u->device = current_device();
    u->dir = direction;
    u->count = count;
    u->flags = flags;
    u->sizes.assign(sizes, sizes + count);
    u->offsets.resize(count);
    long acc = 0;
    for (int i = 0; i < count; ++i) {
        const long n = sizes[i];
        if (n < 2 || (n & (n - 1)) || n > (1L << kUMaxLog2)) {
            ubatch_free(u);
            return set_error(ESTL_ERR_INVALID, "transform %d: size %ld is not a power of two in [2, 2^%d]", i, n,
                             kUMaxLog2);
        }
        u->offsets[i] = offsets ? offsets[i] : acc;
        if (u->offsets[i] < 0) {
            ubatch_free(u);
            return set_error(ESTL_ERR_INVALID, "transform %d: negative offset", i);
        }
        acc += n;
        u->total_elems += n;
        u->extent = std::max(u->extent, u->offsets[i] + n);
    }
    // the normalisation factor is per transform; the kernel applies one scale, so a
    // normalised uneven batch needs one size
    if (flags & ESTL_FLAG_NORMALIZE) {
        for (int i = 1; i < count; ++i)
            if (sizes[i] != sizes[0]) {
                ubatch_free(u);
                return set_error(ESTL_ERR_UNSUPPORTED, "ESTL_FLAG_NORMALIZE needs equal sizes in an uneven batch");
            }
        u->scale = 1.0f / (float)sizes[0];
    }
    // sort by size, largest first (stable: equal sizes keep caller order)
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sizes[a] > sizes[b]; });
    std::vector<long> soffs(count);
    for (int k = 0; k < count; ++k) soffs[k] = u->offsets[order[k]];
    for (int k = 0; k < count;) {
        const int L = ilog2l(sizes[order[k]]);
        int per = (L > kUMaxSmallLog2) ? 1 : ubatch_small_rows_per_item(L);
        int m = 0;
        while (m < per && k + m < count && sizes[order[k + m]] == sizes[order[k]]) ++m;
        UItem it{L, m, k, 0};
        u->items.push_back(it);
        if (L > kUMaxSmallLog2) {
            const size_t end = (size_t)u->offsets[order[k]] + (size_t)sizes[order[k]];
            u->scratch_elems = std::max(u->scratch_elems, end);
            u->large_elems += sizes[order[k]];
        }
        k += m;
    }
    // tables
    UTables &tw = u->args.tw;
    memset(&tw, 0, sizeof tw);
    for (int L = 1; L <= kUMaxSmallLog2; ++L) {
        int rc = power_table(u->device, L, &tw.pow[L]);
        if (rc) {
            ubatch_free(u);
            return rc;
        }
    }
    for (int L = kUMaxSmallLog2 + 1; L <= kUMaxLog2; ++L) {
        bool used = false;
        for (auto &it : u->items) used |= (it.log2n == L);
        if (!used) continue;
        const long nq = 1L << L;
        tw.lobits[L] = (L + 1) / 2;
        int rc = linear_table(u->device, 1L << tw.lobits[L], 1, nq, &tw.lo[L]);
        if (!rc) rc = linear_table(u->device, nq >> tw.lobits[L], 1L << tw.lobits[L], nq, &tw.hi[L]);
        if (rc) {
            ubatch_free(u);
            return rc;
        }
    }
    HIP_TRY_U(hipMalloc(&u->d_items, sizeof(UItem) * u->items.size()));
    HIP_TRY_U(hipMemcpy(u->d_items, u->items.data(), sizeof(UItem) * u->items.size(), hipMemcpyHostToDevice));
    HIP_TRY_U(hipMalloc(&u->d_offs, sizeof(long) * count));
    HIP_TRY_U(hipMemcpy(u->d_offs, soffs.data(), sizeof(long) * count, hipMemcpyHostToDevice));
    HIP_TRY_U(hipMalloc(&u->d_counters, 2 * sizeof(int)));
    HIP_TRY_U(hipMemset(u->d_counters, 0, 2 * sizeof(int)));
    if (u->scratch_elems) HIP_TRY_U(hipMalloc(&u->scratch, sizeof(float2) * u->scratch_elems));
    HIP_TRY_U(hipDeviceSynchronize());
    u->blocks = std::min<long>(ubatch_resident_blocks(), (long)u->items.size());
    u->args.items = u->d_items;
    u->args.n_items = (int)u->items.size();
    u->args.offs = u->d_offs;
    u->args.scratch = u->scratch;
    u->args.counters = u->d_counters;
    u->args.scale = u->scale;
    *out = u;
    return ESTL_OK;
}

static int ubatch_execute(UBatch *u, const void *in, void *out, hipStream_t st) {
    if (!in || !out) return set_error(ESTL_ERR_INVALID, "null buffer");
    UArgs a = u->args;
    a.in = static_cast<const float2 *>(in);
    a.out = static_cast<float2 *>(out);
    hipError_t e = launch_ubatch(u->dir, a, u->blocks, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "k_ubatch launch failed: %s", hipGetErrorString(e));
    return ESTL_OK;
}

// The graph is captured on a private stream (the caller's may be the NULL stream, which
// cannot be captured) and launched on the caller's stream.
static int ubatch_execute_graph(UBatch *u, const void *in, void *out, hipStream_t st) {
    if (!u->gexec || u->g_in != in || u->g_out != out) {
        if (u->gexec) (void)hipGraphExecDestroy(u->gexec);
        if (u->graph) (void)hipGraphDestroy(u->graph);
        u->gexec = nullptr;
        u->graph = nullptr;
        hipError_t e = hipSuccess;
        if (!u->cap_stream) e = hipStreamCreateWithFlags(&u->cap_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "capture stream: %s", hipGetErrorString(e));
        e = hipStreamBeginCapture(u->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "begin capture failed: %s", hipGetErrorString(e));
        int rc = ubatch_execute(u, in, out, u->cap_stream);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(u->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "end capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&u->gexec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(ESTL_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        }
        u->graph = g;
        u->g_in = in;
        u->g_out = out;
    }
    hipError_t e = hipGraphLaunch(u->gexec, st);
    if (e != hipSuccess) return set_error(ESTL_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e));
    return ESTL_OK;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_ubatch_create(estl_ubatch_t *ub, int count, const long *sizes, const long *offsets, int direction,
                       unsigned flags) {
    if (!ub) return set_error(ESTL_ERR_INVALID, "handle pointer is null");
    *ub = nullptr;
    UBatch *u = nullptr;
    int rc = ubatch_create(&u, count, sizes, offsets, direction, flags);
    if (rc) return rc;
    *ub = reinterpret_cast<estl_ubatch_t>(u);
    return ESTL_OK;
}

int estl_ubatch_execute(estl_ubatch_t ub, const void *d_in, void *d_out, void *stream) {
    if (!ub) return set_error(ESTL_ERR_INVALID, "handle is null");
    return ubatch_execute(reinterpret_cast<UBatch *>(ub), d_in, d_out, static_cast<hipStream_t>(stream));
}

int estl_ubatch_execute_graph(estl_ubatch_t ub, const void *d_in, void *d_out, void *stream) {
    if (!ub) return set_error(ESTL_ERR_INVALID, "handle is null");
    return ubatch_execute_graph(reinterpret_cast<UBatch *>(ub), d_in, d_out, static_cast<hipStream_t>(stream));
}

int estl_ubatch_describe(estl_ubatch_t ub, char *buf, int len) {
    if (!ub) return set_error(ESTL_ERR_INVALID, "handle is null");
    UBatch *u = reinterpret_cast<UBatch *>(ub);
    int small_items = 0, large_items = 0;
    for (auto &it : u->items) (it.log2n > kUMaxSmallLog2 ? large_items : small_items)++;
    char tmp[512];
    snprintf(tmp, sizeof tmp,
             "{\"transforms\": %d, \"items\": %zu, \"small_items\": %d, \"large_items\": %d, "
             "\"workgroups\": %ld, \"elements\": %ld, \"extent\": %ld, \"two_pass_elements\": %ld, "
             "\"launches\": 1}",
             u->count, u->items.size(), small_items, large_items, u->blocks, u->total_elems, u->extent,
             u->large_elems);
    if (buf && len > 0) {
        strncpy(buf, tmp, (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)strlen(tmp);
}

int estl_ubatch_destroy(estl_ubatch_t ub) {
    ubatch_free(reinterpret_cast<UBatch *>(ub));
    return ESTL_OK;
}

}  // extern "C"
