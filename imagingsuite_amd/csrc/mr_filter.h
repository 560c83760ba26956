// Row ramp-filter launch interface shared by the parallel-beam and FDK paths.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace mr {

// FullLogNorm of one pixel (NormPlugins.cpp:411-558, non-LUT path):
// p = img - dark (or img); p <= 0 ? 0 : flat_log - log(p) + dl (no flat:
// -log(p) + dl).  Shared by the normalisation kernel and the filter's fused
// load, so that both give the same values.
__device__ __forceinline__ float lognorm_px(float v, const float *flat_log, const float *dark, size_t i, float dl)
{
    const float p = dark ? v - dark[i] : v;
    if (p <= 0.0f) return 0.0f;
    return flat_log ? flat_log[i] - logf(p) + dl : -logf(p) + dl;
}

struct RowFilterArgs {
    float *rows;                // n_rows x n, filtered in place
    long n_rows;
    int n;                      // row width
    int mode;                   // 0 ProjectionFilterSingle, 1 FDK ramp_filter_tuned,
                                // 2 adjoint of mode 0, 3 ProjectionFilterSingle as a
                                // short convolution (see resp_first), 4 adjoint of
                                // mode 3
    int insert;                 // mode 0: nInsert; mode 1: (P - n) / 2
    const float *taper;         // mode 0: nInsert sigmoid weights
    const float *hs;            // P real spectral multipliers
    const float2 *tw;           // P twiddles exp(-2 pi i k / P)
    float post0, post1, post2;  // sequential output multipliers
    const float *pre_w;         // mode 1: per-projection input weight (nullable)
    int rows_per_proj;          // rows of one projection (mode 1)
    int cosw;                   // mode 1: multiply by the Feldkamp cosine weight
    float cw_u, cw_v;           // pix_u / SDD, pix_v / SDD
    float cw_c0, cw_r0;         // piercing point (column, row) in the window
    float *out;                 // destination rows (nullptr: in place)
    // mode 3: the row less the straight line from its first to its last
    // value (s = i * ramp_inv, line = first (1 - s) + last s) is convolved
    // through the short spectrum in hs; resp_first / resp_last are the
    // filter's responses at the n output samples to the left taper plus
    // 1 - s on the row, and to the right taper plus s on the row.  Mode 4:
    // the responses to the two tapers alone, whose dot products with the
    // row fold back onto its first and last sample
    const float *resp_first, *resp_last;
    float ramp_inv;             // 1 / (n - 1) in float
    // mode 3 with norm set: each row is FullLogNorm-normalised on load (row
    // q of projection j: pixels q n .. q n + n - 1 of flat_log / dark, dose
    // term dl[j]; rows_per_proj rows per projection)
    int norm;
    const float *norm_flat_log, *norm_dark, *norm_dl;
};

int run_row_filter(const RowFilterArgs &a, int P, hipStream_t s);
}  // namespace mr

struct mr_filter_plan;

namespace mr {
// ProjectionFilterSingle from `in` into `out` (out may equal in).
int filter_rows_to(mr_filter_plan *plan, const float *in, float *out, size_t n_rows, hipStream_t s);
size_t filter_width(const mr_filter_plan *plan);
int run_fdk_margin_rows(float *img, int n, int h, long n_proj, hipStream_t s);
std::vector<float> make_twiddles(size_t P);
// Zero-fills n floats with a kernel (no hipMemset, so callers may capture it in a graph).
int fill_zero(float *d, size_t n, hipStream_t s);

}  // namespace mr
