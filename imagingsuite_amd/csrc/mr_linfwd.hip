// LinearForwardProjector::project on gfx950
// (frameworks/tomography/Framework/ReconAlgorithms/ReconAlgorithms/
// linearforwardprojector.cpp:15-143), for a stack of slices instead of one.
//
// For every angle t and bin r the reference steps along the axis the ray is
// closest to (rows m when |sin| > sqrt(1/2), else columns n), keeps the
// crossing position in a float that it advances by alpha per step,
// truncates it, and sums slice(p) + w (slice(p+1) - slice(p)) while the
// position is in [0, N-1); the sum is scaled by 1/|sin| (or 1/|cos|) in
// double.  One thread follows one (angle, bin) ray for up to four slices
// with exactly those float operations, so each slice's projection is the
// reference's bit for bit.  Per-angle constants (float trig of the float
// angle, the double step factors) come from the host in the reference's
// types.
//
// Layout: volume [M(y)][M(x)][nz] (MatrixZXY; slice(x, y) of slice z is
// vol[(y M + x) nz + z]), sinogram [n_ang][M][nz] (proj(r, t) of slice z).
#include "mr_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mr {

struct LinFwdAngle {
    float alpha, rhooffset;
    int rows;         // 1: step over m with slice(m, nf) (|sin| > sqrt(1/2))
    int pad;
    double dx, dxd;   // Delta_x sin (or cos), Delta_x / |sin| (or |cos|)
};

__global__ __launch_bounds__(256) void linear_forward_kernel(const float *__restrict__ vol, float *__restrict__ sino,
                                                             const LinFwdAngle *__restrict__ tab, int M, int nz,
                                                             int zq_per_block)
{
    const int rays = 256 / zq_per_block;
    const int r = blockIdx.x * rays + (int)threadIdx.x % rays;
    const int t = blockIdx.y;
    const int z0 = 4 * (blockIdx.z * zq_per_block + (int)threadIdx.x / rays);
    if (r >= M || z0 >= nz) return;
    const int nzl = min(4, nz - z0);
    const LinFwdAngle a = tab[t];
    const float eps = 1e-4f;
    const float rho_min = (float)(-M / 2);
    const float rho = (float)r;
    const float beta = (float)((double)(rho + rho_min - a.rhooffset) / a.dx);
    const float betap = (float)((double)beta + 0.5);
    const float alpha = a.alpha;
    int lo, hi;
    if (alpha > 1e-6f) {
        lo = (int)ceilf(-(betap - eps) / alpha);
        hi = 1 + (int)floorf(((float)M - betap - eps) / alpha);
    } else if (alpha < -1e-6f) {
        lo = (int)ceilf(((float)M - betap - eps) / alpha);
        hi = 1 + (int)floorf(-(betap - eps) / alpha);
    } else {
        lo = 0;
        hi = M;
    }
    if (lo < 0) lo = 0;
    if (lo >= M) lo = M - 1;
    if (hi > M) hi = M;
    float pos = beta + (float)lo * alpha;
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const float lim = (float)(M - 1);
    const bool vec = nzl == 4 && (nz & 3) == 0;
    for (int k = lo; k < hi; ++k) {
        const float pf = (float)(int)pos;
        const float w = pos - pf;
        if (pos >= 0.0f && pos < lim) {
            const int p = (int)pf;
            // rows: slice(m = k, nf = p) and slice(k, p + 1); columns: slice(mf = p, n = k), slice(p + 1, k)
            const size_t ia = a.rows ? (size_t)p * M + k : (size_t)k * M + p;
            const size_t ib = a.rows ? ia + (size_t)M : ia + 1;
            const float *pa = vol + ia * nz + z0, *pb = vol + ib * nz + z0;
            if (vec) {
                const float4 va = *reinterpret_cast<const float4 *>(pa), vb = *reinterpret_cast<const float4 *>(pb);
                sum[0] += va.x + w * (vb.x - va.x);
                sum[1] += va.y + w * (vb.y - va.y);
                sum[2] += va.z + w * (vb.z - va.z);
                sum[3] += va.w + w * (vb.w - va.w);
            } else {
                for (int j = 0; j < nzl; ++j) sum[j] += pa[j] + w * (pb[j] - pa[j]);
            }
        }
        pos += alpha;
    }
    float *out = sino + ((size_t)t * M + r) * nz + z0;
    for (int j = 0; j < nzl; ++j) out[j] = (float)((double)sum[j] * a.dxd);
}

// Per-angle constants with the reference's types and operation order
// (float angle in radians, std::sin/std::cos of a float, double step factors).
std::vector<LinFwdAngle> linfwd_tables(const float *angles, size_t n_ang, int M)
{
    const float fpi = 3.1415926535897932384626f;
    const float x_min = (float)(-M / 2), Delta_x = 1.0f;
    std::vector<LinFwdAngle> tab(n_ang);
    for (size_t t = 0; t < n_ang; ++t) {
        float theta = angles[t] + 0.0f;
        theta *= fpi / 180.0f;
        const float st = std::sin(theta), ct = std::cos(theta);
        LinFwdAngle &e = tab[t];
        e.rhooffset = x_min * (ct + st);
        e.pad = 0;
        if ((double)std::fabs(st) > std::sqrt(0.5)) {
            e.rows = 1;
            e.alpha = -ct / st;
            e.dx = (double)(Delta_x * st);
            e.dxd = (double)(Delta_x / std::fabs(st));
        } else {
            e.rows = 0;
            e.alpha = -st / ct;
            e.dx = (double)(Delta_x * ct);
            e.dxd = (double)(Delta_x / std::fabs(ct));
        }
    }
    return tab;
}

}  // namespace mr

extern "C" int mr_linear_forward(const float *d_volume, size_t M, size_t nz, const float *angles, size_t n_ang,
                                 float *d_sino, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(M >= 2 && M <= (1u << 20), "slice size out of range");
    MR_REQUIRE(nz >= 1, "at least one slice");
    MR_REQUIRE(n_ang >= 1 && n_ang <= 65535, "angle count out of range (1..65535)");
    MR_REQUIRE(d_volume != nullptr && d_sino != nullptr && angles != nullptr, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    std::vector<mr::LinFwdAngle> tab = mr::linfwd_tables(angles, n_ang, (int)M);
    mr::LinFwdAngle *d_tab = nullptr;
    MR_HIP(hipMallocAsync((void **)&d_tab, sizeof(mr::LinFwdAngle) * n_ang, s));
    hipError_t e = hipMemcpyAsync(d_tab, tab.data(), sizeof(mr::LinFwdAngle) * n_ang, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const size_t ngroups = (nz + 3) / 4;
        const int zqb = ngroups >= 4 ? 4 : (ngroups >= 2 ? 2 : 1);
        const int rays = 256 / zqb;
        dim3 grid((unsigned)((M + rays - 1) / rays), (unsigned)n_ang, (unsigned)((ngroups + zqb - 1) / zqb));
        mr::ProfScope prof("linear_forward", s);
        hipLaunchKernelGGL(mr::linear_forward_kernel, grid, dim3(256), 0, s, d_volume, d_sino, d_tab, (int)M,
                           (int)nz, zqb);
        e = hipGetLastError();
    }
    hipError_t e2 = hipFreeAsync(d_tab, s);
    if (e != hipSuccess) return mr::hip_fail(e, "linear_forward_kernel", __FILE__, __LINE__);
    if (e2 != hipSuccess) return mr::hip_fail(e2, "hipFreeAsync", __FILE__, __LINE__);
    // the host table must outlive the copy
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
    MR_GUARD_END
}
