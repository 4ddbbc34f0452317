// ols_model_test.cpp -- the overlap-save window choice (conv_ols_log2, conv_ols_both in
// estl_amd/csrc/plan.h) on the host, against the lengths that measured clearly faster
// (by 5 % or more) on an MI355X of 256 CUs, and the bounds of each path.
// Usage: ols_model_test; prints PASS/FAIL lines, exits 0 when all pass.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../estl_amd/csrc/plan.h"

using namespace estl;

static int g_failures = 0;

static void expect(long n, long rows, long k, unsigned flags, int want) {
    const int got = conv_ols_log2(n, rows, k, flags, 256);
    const bool ok = got == want;
    if (!ok) ++g_failures;
    printf("%s n=2^%d rows=%ld k=%ld flags=0x%x: log2 window %d, want %d (0: whole rows)\n", ok ? "PASS" : "FAIL",
           __builtin_ctzl((unsigned long)n), rows, k, flags, got, want);
}

int main() {
    // measured: the faster window length where the two differ by 5 % or more
    expect(1L << 22, 64, 16, 0, 13);
    expect(1L << 22, 64, 3000, 0, 14);
    expect(1L << 22, 4, 16, 0, 13);
    expect(1L << 22, 4, 2048, 0, 13);
    expect(1L << 22, 4, 4097, 0, 14);
    expect(1L << 21, 16, 1536, 0, 13);
    expect(1L << 21, 16, 4097, 0, 14);
    expect(1L << 20, 256, 1024, 0, 13);
    expect(1L << 20, 256, 3000, 0, 14);
    expect(1L << 20, 8, 2048, 0, 13);
    expect(1L << 16, 64, 2048, 0, 14);
    expect(1L << 19, 24, 1024, 0, 13);
    expect(1L << 18, 64, 2048, 0, 13);
    expect(1L << 18, 64, 4097, 0, 14);
    expect(1L << 16, 4096, 3000, 0, 14);
    expect(1L << 16, 256, 4097, 0, 14);
    expect(1L << 15, 8192, 2048, 0, 13);
    // bounds: only one length fits, or neither
    expect(1L << 14, 3, 4097, 0, 13);   // rows of 2^14: no 2^14 windows
    expect(1L << 14, 3, 4098, 0, 0);    // ... and no 2^13 window holds 4098 taps
    expect(1L << 15, 3, 4098, 0, 14);   // above 4097 taps only 2^14 windows
    expect(1L << 22, 1, 8193, 0, 14);
    expect(1L << 22, 1, 8194, 0, 0);    // beyond 8193 taps: whole rows
    expect(1L << 13, 5, 16, 0, 0);      // a row fits one launch of the fused kernel
    expect(1L << 20, 8, 16, ESTL_CONV_FULL_LENGTH, 0);
    expect(1L << 20, 8, 16, ESTL_CONV_CORRELATE | ESTL_CONV_CENTER, 13);  // other flags do not matter
    expect(1L << 20, 0, 16, 0, 0);      // no rows
    expect(1L << 20, 8, 0, 0, 0);       // no taps
    // window geometry: the first valid position and the outputs per window are multiples of
    // the alignment (32 FP32 / 64 binary16 values), the valid outputs stay inside the part
    // of the window where the L-point circular result equals the n-point one, and at most
    // 2 x alignment - 1 outputs of L - K + 1 are given up
    {
        long bad = 0, checked = 0;
        for (int lw = 13; lw <= 14; ++lw)
            for (long k = 1; k <= (lw == 13 ? 4097 : 8193); ++k)
                for (int corr = 0; corr < 2; ++corr)
                    for (int h = 0; h < 2; ++h) {
                        const long L = 1L << lw, al = ols_align(h), pre = corr ? 0 : k - 1;
                        const long f = ols_first_valid(k, corr, h), m = ols_outputs(lw, k, corr, h);
                        const long end = corr ? L - k + 1 : L;  // valid window positions [pre, end)
                        ++checked;
                        if (f % al || m % al || f < pre || f + m > end || m < 1 || m < (L - k + 1) - (2 * al - 1))
                            ++bad;
                    }
        if (bad) ++g_failures;
        printf("%s window geometry: %ld of %ld (length, taps, correlation, precision) cases off\n",
               bad ? "FAIL" : "PASS", bad, checked);
    }
    // a registration keeps both spectra exactly where both lengths are possible
    const bool both_ok = conv_ols_both(1L << 15, 4097, 0) && !conv_ols_both(1L << 15, 4098, 0) &&
                         !conv_ols_both(1L << 14, 100, 0) && !conv_ols_both(1L << 16, 100, ESTL_CONV_FULL_LENGTH) &&
                         conv_ols_wide_offset(false) == 8192 && conv_ols_wide_offset(true) == 8193;
    if (!both_ok) ++g_failures;
    printf("%s conv_ols_both / conv_ols_wide_offset\n", both_ok ? "PASS" : "FAIL");
    printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
    return g_failures ? 1 : 0;
}
