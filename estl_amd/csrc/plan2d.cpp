// plan2d.cpp -- 2-D transforms on the device: the role of FFT2D (FFT.h:54-84), which
// runs one 1-D execute (one host round trip) per row and then per column.
//
// Layout: `batch` arrays of ny rows of nx complex64, x fastest (FFT2D's data[y*xDim + x]).
//   rows:    the 1-D plan of length nx over batch*ny rows (one launch for nx <= 2^13);
//   columns: no transpose.  A column transform of length ny at stride nx is the ny plan's
//            Stockham passes with every index scaled by nx: in pass_tile's terms the
//            "row" is the whole nx*ny array (log2n + log2 nx), the stride S grows by nx
//            (log2S + log2 nx), and A_q, B_q and the pre-twiddles W_{A B}^{g s} stay
//            those of the ny transform.  For ny <= 2^13 that is one pass of radix ny, in
//            place (every tile writes exactly the positions it read).
// So a 2-D transform makes two HBM round trips for sizes up to 2^13 x 2^13.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/estl_hip.h"
#include "internal.h"
#include "plan.h"

namespace estl {

int default_pass_tile(int log2b, int log2cpr, long total_cols);

struct Plan2D {
    long nx = 0, ny = 0, batch = 0;
    int dir = ESTL_FORWARD;
    unsigned flags = 0;
    Plan rows, cols;
    bool has_rows = false, has_cols = false;
    float2 *h_in_dev = nullptr, *h_out_dev = nullptr;
    hipStream_t stream = nullptr;
};

static int ilog2p(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

static int build_plan2d(Plan2D *p, long nx, long ny, long batch, int dir, unsigned flags) {
    if (nx < 1 || (nx & (nx - 1)) || ny < 1 || (ny & (ny - 1)))
        return set_error(ESTL_ERR_INVALID, "%ld x %ld: both sizes must be powers of two", nx, ny);
    if (batch < 1) return set_error(ESTL_ERR_INVALID, "batch=%ld must be >= 1", batch);
    if (dir != ESTL_FORWARD && dir != ESTL_BACKWARD) return set_error(ESTL_ERR_INVALID, "direction=%d", dir);
    if (flags & ~(unsigned)ESTL_FLAG_NORMALIZE) return set_error(ESTL_ERR_INVALID, "unsupported flags 0x%x", flags);
    const int lx = ilog2p(nx), ly = ilog2p(ny);
    if (lx + ly > 30) return set_error(ESTL_ERR_UNSUPPORTED, "nx*ny = 2^%d exceeds 2^30", lx + ly);
    p->nx = nx;
    p->ny = ny;
    p->batch = batch;
    p->dir = dir;
    p->flags = flags;
    const float scale = (flags & ESTL_FLAG_NORMALIZE) ? 1.0f / ((float)nx * (float)ny) : 1.0f;
    // the last launch applies the normalisation
    p->has_cols = ny > 1;
    p->has_rows = nx > 1 || !p->has_cols;
    if (p->has_rows) {
        int rc = build_plan(&p->rows, nx, batch * ny, dir, ESTL_PRECISION_FP32,
                            p->has_cols ? 0u : (flags & ESTL_FLAG_NORMALIZE), nullptr);
        if (rc) return rc;
    }
    if (p->has_cols) {
        // Up to ny = 2^9 the columns take one pass of radix ny, whose tiles keep 2^(13 -
        // log2 ny) >= 16 adjacent columns (128-byte runs).  A longer column keeps fewer
        // (two at 4096: 16-byte runs at a 32 KiB stride), so ny = 2^10..2^13 takes two
        // passes of radix <= 2^7 with 64-column tiles (512-byte runs): one more HBM round
        // trip, measured much faster.  Above 2^13 the default split applies.
        std::vector<int> split;
        if (ly <= 9 || (ly <= kMaxRowsLog2 && nx < 4)) {
            split.push_back(ly);
        } else if (ly <= kMaxRowsLog2) {
            split.push_back(ly / 2);
            split.push_back(ly - ly / 2);
        }
        int rc = build_plan(&p->cols, ny, 1, dir, ESTL_PRECISION_FP32, 0, split.empty() ? nullptr : &split);
        if (rc) return rc;
        // rescale the ny plan's passes to columns of stride nx
        for (size_t q = 0; q < p->cols.steps.size(); ++q) {
            PassStep &s = p->cols.steps[q];
            s.args.log2n += lx;
            s.args.log2S += lx;
            s.args.batch = batch;
            s.args.xcd_remap = 0;
            s.args.scale = (q + 1 == p->cols.steps.size()) ? scale : 1.0f;
            s.log2t = default_pass_tile(s.log2b, s.args.log2n - s.log2b, batch * ((nx * ny) >> s.log2b));
        }
        p->cols.n = nx * ny;
        p->cols.batch = batch;
        p->cols.log2n = lx + ly;
        p->cols.scratch_bytes = sizeof(float2) * (size_t)nx * (size_t)ny * (size_t)batch;
    }
    return ESTL_OK;
}

static int execute_plan2d(Plan2D *p, const float2 *in, float2 *out, hipStream_t st) {
    if (!in || !out) return set_error(ESTL_ERR_INVALID, "null buffer");
    const float2 *src = in;
    if (p->has_rows) {
        int rc = execute_plan(&p->rows, in, out, st);
        if (rc) return rc;
        src = out;
    }
    if (p->has_cols) return execute_plan(&p->cols, src, out, st);
    return ESTL_OK;
}

static void destroy_plan2d(Plan2D *p) {
    if (!p) return;
    for (Plan *q : {&p->rows, &p->cols}) {
        free_stream_scratch(q);
        for (auto &s : q->scratch)
            if (s) (void)hipFree(s);
    }
    if (p->h_in_dev) (void)hipFree(p->h_in_dev);
    if (p->h_out_dev) (void)hipFree(p->h_out_dev);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

}  // namespace estl

using namespace estl;

extern "C" {

int estl_plan2d_create(estl_plan2d_t *plan, long nx, long ny, long batch, int direction, unsigned flags) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan pointer is null");
    *plan = nullptr;
    Plan2D *p = new Plan2D();
    int rc = build_plan2d(p, nx, ny, batch, direction, flags);
    if (rc) {
        destroy_plan2d(p);
        return rc;
    }
    *plan = reinterpret_cast<estl_plan2d_t>(p);
    return ESTL_OK;
}

int estl_plan2d_execute(estl_plan2d_t plan, const void *d_in, void *d_out, void *stream) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    return execute_plan2d(reinterpret_cast<Plan2D *>(plan), static_cast<const float2 *>(d_in),
                          static_cast<float2 *>(d_out), static_cast<hipStream_t>(stream));
}

int estl_plan2d_execute_host(estl_plan2d_t plan, const void *h_in, void *h_out) {
    if (!plan || !h_in || !h_out) return set_error(ESTL_ERR_INVALID, "null argument");
    Plan2D *p = reinterpret_cast<Plan2D *>(plan);
    const size_t bytes = sizeof(float2) * (size_t)p->nx * (size_t)p->ny * (size_t)p->batch;
    if (!p->stream) HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    if (!p->h_in_dev) HIP_TRY(hipMalloc(&p->h_in_dev, bytes));
    if (!p->h_out_dev) HIP_TRY(hipMalloc(&p->h_out_dev, bytes));
    HIP_TRY(hipMemcpyAsync(p->h_in_dev, h_in, bytes, hipMemcpyHostToDevice, p->stream));
    int rc = execute_plan2d(p, p->h_in_dev, p->h_out_dev, p->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, p->h_out_dev, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return ESTL_OK;
}

int estl_plan2d_launches(estl_plan2d_t plan) {
    if (!plan) return set_error(ESTL_ERR_INVALID, "plan is null");
    const Plan2D *p = reinterpret_cast<Plan2D *>(plan);
    return (p->has_rows ? plan_launches(&p->rows) : 0) + (p->has_cols ? plan_launches(&p->cols) : 0);
}

int estl_plan2d_destroy(estl_plan2d_t plan) {
    destroy_plan2d(reinterpret_cast<Plan2D *>(plan));
    return ESTL_OK;
}

}  // extern "C"
