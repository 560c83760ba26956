// FDK cone-beam back-projection on gfx950 (FDKbp_single).
//
// Data layout in HBM
//   d_proj  [n_proj][win_rows][width]  window projections; weighted, filtered and
//                                      scaled in place (fdkbp_single.cpp:376-389)
//   volume  [nz][ny][nx]               MatrixXYZ, written already mirrored in x and
//                                      flipped in z (fdkreconbase.cpp:259-270)
//
// fdk_kernel: one thread owns one (x, y) voxel column and KT consecutive
// slices in registers.  With the reference geometry (detector up = z, no
// tilt) the homogeneous depth and the detector column do not depend on z, so
// 1/depth, its square and the column index are computed once per projection
// and column; only the detector row changes along z.  The per-projection
// coefficients are the reference's xip/yip/zip products
// (fdkbp_single.cpp:420-456) evaluated on the host in float, and the
// accumulation order per voxel is the reference's projection order, so the
// result is bit-identical to FDKbp_single.
#include "mr_common.h"
#include "mr_host.h"
#include "mr_filter.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

namespace mr {

struct FdkArgs {
    const float *img;
    float *vol;
    const float *ptab;  // 8 floats per projection: cx0 cx1 cx2 cy0 cy1 cy2 zc0 zc2
    const float *zt1;   // [np][nz] zip row term per projection and slice
    const float *zt0;   // [np][nz] zip column term per slice (tilt correction only)
    const int2 *mask;   // (first, second]: reference loop i = first+1 .. second
    int nx, ny, nz, w, h, np;
    float ox, oy, sx, sy;
    int order, ntx, nty, ntz;
    int wx;             // voxel columns along x per wave (a power of two up to 64)
    int bwx;            // waves along x per block (1, 2 or 4)
};

// Thread -> voxel column.  A wave covers wx x (64 / wx) columns and a block
// bwx x (4 / bwx) waves.  A compact wave footprint keeps a gather's 64
// samples on fewer detector rows (the row of slice k moves with the depth,
// through 1/depth, which changes along the ray direction), and a compact
// block footprint lets its waves share the detector lines they fetch.
__device__ __forceinline__ void fdk_voxel(int wx, int bwx, int bx, int by, int &i, int &j)
{
    const int l = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int wy = 64 / wx, bwy = 4 / bwx;
    i = bx * bwx * wx + (w & (bwx - 1)) * wx + (l & (wx - 1));
    j = by * bwy * wy + (w / bwx) * wy + l / wx;
}

// Block -> tile mapping: order 0 is the 3-D grid; order 1 a 1-D grid whose
// consecutive blocks go to consecutive XCDs (round robin over 8), each XCD
// taking a contiguous eighth of the tiles in slice-chunk, row, column order,
// so the blocks resident on one XCD share detector rows in its L2.
__device__ __forceinline__ bool fdk_tile(int order, int ntx, int nty, int ntz, int &bx, int &by, int &bz)
{
    if (order == 0) {
        bx = blockIdx.x;
        by = blockIdx.y;
        bz = blockIdx.z;
        return true;
    }
    const int n = ntx * nty * ntz;
    const int per = (n + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if (t >= n) return false;
    bz = t / (ntx * nty);
    const int r = t - bz * ntx * nty;
    by = r / ntx;
    bx = r - by * ntx;
    return true;
}

// Detector samples are read with buffer loads whose range is one projection
// (h * w floats): the row index is clamped to [-1, h] and floored, so rows
// outside the window (r < 0 or r >= h, the reference's two tests) give byte
// offsets outside the range, which the hardware returns as 0.0 - the value
// get_pixel_value_c returns there.  Inside the window floor(r) = (int)r.
// The KT row terms are fetched first, then all KT samples are requested
// before any of them is used, so the loads of one projection are in flight
// together.
// (int)floor(v) in one instruction for finite v (the clamped row position):
// the compiler keeps floor and convert apart because they differ for NaN.
__device__ __forceinline__ int cvt_floor(float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
#else
    return (int)floorf(v);   // host pass of a device-only function: never executed
#endif
}

template <int KT, bool FULL, bool FUSED>
__device__ __forceinline__ void fdk_column(float (&acc)[KT], const float *__restrict__ zt, int kn,
                                           __amdgpu_buffer_rsrc_t img, float yip1, float xip1, float dw,
                                           float dw2, float fh, int w4, int cc4)
{
    float ztv[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) ztv[kk] = (FULL || kk < kn) ? zt[kk] : 0.0f;
    // fused: r = dw * zt + (dw * (yip1 + xip1) + 0.5), one FMA per slice
    const float rb = FUSED ? __builtin_fmaf(dw, yip1 + xip1, 0.5f) : 0.0f;
    float val[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
        float r;
        if (FUSED) {
            r = __builtin_fmaf(dw, ztv[kk], rb);
        } else {
            const float acc21 = ztv[kk] + yip1;
            r = dw * (acc21 + xip1);
            r = r + 0.5f;
        }
        const int rr = cvt_floor(__builtin_amdgcn_fmed3f(r, -1.0f, fh));
        val[kk] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(img, __mul24(rr, w4) + cc4, 0, 0));
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (FULL || kk < kn) acc[kk] = FUSED ? __builtin_fmaf(dw2, val[kk], acc[kk]) : acc[kk] + dw2 * val[kk];
}

// FUSED = false: the reference's float operations (bit-identical positions
// and sums).  FUSED = true: the row position folded to one FMA per slice and
// the accumulation to an FMA (positions can round differently at pixel
// edges, where nearest-neighbour sampling picks the neighbouring pixel).
template <int KT, bool FUSED>
__global__ __launch_bounds__(256) void fdk_kernel(FdkArgs a)
{
    int bx, by, bz;
    if (!fdk_tile(a.order, a.ntx, a.nty, a.ntz, bx, by, bz)) return;
    int i, j;
    fdk_voxel(a.wx, a.bwx, bx, by, i, j);
    const int k0 = bz * KT;
    if (i >= a.nx || j >= a.ny) return;
    const int kn = min(KT, a.nz - k0);
    float acc[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0f;
    const int2 m = a.mask[j];
    if (i > m.x && i <= m.y) {
        const float x = a.ox + (float)i * a.sx;
        const float y = a.oy + (float)j * a.sy;
        const float fw = (float)a.w, fh = (float)a.h;
        const int w4 = 4 * a.w;
        const int pbytes = 4 * a.w * a.h;
        for (int p = 0; p < a.np; ++p) {
            const float *t = a.ptab + 8 * p;
            const float xip0 = x * t[0], xip1 = x * t[1], xip2 = x * t[2];
            const float yip0 = y * t[3], yip1 = y * t[4], yip2 = y * t[5];
            const float acc20 = t[6] + yip0;
            const float acc22 = t[7] + yip2;
            const float dw = 1.0f / (acc22 + xip2);
            float c = dw * (acc20 + xip0);
            c = c + 0.5f;
            if (c < 0.0f || c >= fw) continue;
            const int cc4 = 4 * (int)c;
            const float dw2 = dw * dw;
            const __amdgpu_buffer_rsrc_t img = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(a.img + (size_t)p * a.h * a.w), (short)0, pbytes, 0x00020000);
            const float *zt = a.zt1 + (size_t)p * a.nz + k0;
            if (kn == KT) fdk_column<KT, true, FUSED>(acc, zt, kn, img, yip1, xip1, dw, dw2, fh, w4, cc4);
            else fdk_column<KT, false, FUSED>(acc, zt, kn, img, yip1, xip1, dw, dw2, fh, w4, cc4);
        }
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (kk < kn)
            a.vol[((size_t)(a.nz - 1 - (k0 + kk)) * a.ny + j) * a.nx + (a.nx - 1 - i)] = acc[kk];
}

// FDKbp (fdkbp.cpp:324-510, 612-631): the same loop with the geometry in
// double.  Voxel coordinates are float expressions widened to double
// (origin + i*spacing in float), the per-projection coefficients and the
// per-slice row term are the reference's double xip/yip/zip products
// (evaluated on the host), the depth reciprocal, the sampling positions and
// the bounds tests are double, and the float voxel takes vol + dw*dw*val
// evaluated in double, once per projection.  With -ffp-contract=off and the
// correctly rounded double division every step matches the reference.
struct FdkdArgs {
    const float *img;
    float *vol;
    const double *ptab;  // 8 doubles per projection: cx0 cx1 cx2 cy0 cy1 cy2 zc0 zc2
    const double *zt1;   // [np][nz]
    const double *zt0;   // [np][nz] column term per slice (tilt correction only)
    const int2 *mask;
    int nx, ny, nz, w, h, np;
    float ox, oy, sx, sy;
    int order, ntx, nty, ntz;
    int wx, bwx;
};

template <int KT, bool FULL>
__device__ __forceinline__ void fdkd_column(float (&acc)[KT], const double *__restrict__ zt, int kn,
                                            __amdgpu_buffer_rsrc_t img, double yip1, double xip1, double dw,
                                            double dw2, double fh, int w4, int cc4)
{
    float val[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
        const double acc21 = ((FULL || kk < kn) ? zt[kk] : 0.0) + yip1;
        const double acc31 = acc21 + xip1;
        double r = acc31 * dw;
        r = r + 0.5;
        const int rr = (int)floor(fmin(fmax(r, -1.0), fh));
        val[kk] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(img, __mul24(rr, w4) + cc4, 0, 0));
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (FULL || kk < kn) acc[kk] = (float)((double)acc[kk] + dw2 * (double)val[kk]);
}

template <int KT>
__global__ __launch_bounds__(256) void fdkd_kernel(FdkdArgs a)
{
    int bx, by, bz;
    if (!fdk_tile(a.order, a.ntx, a.nty, a.ntz, bx, by, bz)) return;
    int i, j;
    fdk_voxel(a.wx, a.bwx, bx, by, i, j);
    const int k0 = bz * KT;
    if (i >= a.nx || j >= a.ny) return;
    const int kn = min(KT, a.nz - k0);
    float acc[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0f;
    const int2 m = a.mask[j];
    if (i > m.x && i <= m.y) {
        const double x = (double)(a.ox + (float)i * a.sx);
        const double y = (double)(a.oy + (float)j * a.sy);
        const double fw = (double)a.w, fh = (double)a.h;
        const int w4 = 4 * a.w;
        const int pbytes = 4 * a.w * a.h;
        for (int p = 0; p < a.np; ++p) {
            const double *t = a.ptab + 8 * p;
            const double xip0 = x * t[0], xip1 = x * t[1], xip2 = x * t[2];
            const double yip0 = y * t[3], yip1 = y * t[4], yip2 = y * t[5];
            const double acc20 = t[6] + yip0;
            const double acc22 = t[7] + yip2;
            const double acc30 = acc20 + xip0;
            const double acc32 = acc22 + xip2;
            const double dw = 1.0 / acc32;
            double c = acc30 * dw;
            c = c + 0.5;
            if (c < 0.0 || c >= fw) continue;
            const int cc4 = 4 * (int)c;
            const double dw2 = dw * dw;
            const __amdgpu_buffer_rsrc_t img = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(a.img + (size_t)p * a.h * a.w), (short)0, pbytes, 0x00020000);
            const double *zt = a.zt1 + (size_t)p * a.nz + k0;
            if (kn == KT) fdkd_column<KT, true>(acc, zt, kn, img, yip1, xip1, dw, dw2, fh, w4, cc4);
            else fdkd_column<KT, false>(acc, zt, kn, img, yip1, xip1, dw, dw2, fh, w4, cc4);
        }
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (kk < kn)
            a.vol[((size_t)(a.nz - 1 - (k0 + kk)) * a.ny + j) * a.nx + (a.nx - 1 - i)] = acc[kk];
}

// Axis-tilt correction (fdkbp_single.cpp:443-448, fdkbp.cpp:452-458): the
// centre of rotation of slice k replaces the matrix's column offset, so the
// zip column term, and with it the detector column, depends on the slice.
// These kernels evaluate column and row per sample, in the reference's
// order, with get_pixel_value_c's four bounds tests.
template <int KT>
__global__ __launch_bounds__(256) void fdk_tilt_kernel(FdkArgs a)
{
    int bx, by, bz;
    if (!fdk_tile(a.order, a.ntx, a.nty, a.ntz, bx, by, bz)) return;
    int i, j;
    fdk_voxel(a.wx, a.bwx, bx, by, i, j);
    const int k0 = bz * KT;
    if (i >= a.nx || j >= a.ny) return;
    const int kn = min(KT, a.nz - k0);
    float acc[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0f;
    const int2 m = a.mask[j];
    if (i > m.x && i <= m.y) {
        const float x = a.ox + (float)i * a.sx;
        const float y = a.oy + (float)j * a.sy;
        const float fw = (float)a.w, fh = (float)a.h;
        for (int p = 0; p < a.np; ++p) {
            const float *t = a.ptab + 8 * p;
            const float xip0 = x * t[0], xip1 = x * t[1], xip2 = x * t[2];
            const float yip0 = y * t[3], yip1 = y * t[4], yip2 = y * t[5];
            const float acc22 = t[7] + yip2;
            const float dw = 1.0f / (acc22 + xip2);
            const float dw2 = dw * dw;
            const float *img = a.img + (size_t)p * a.h * a.w;
            const float *z0t = a.zt0 + (size_t)p * a.nz + k0;
            const float *z1t = a.zt1 + (size_t)p * a.nz + k0;
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) {
                if (kk < kn) {
                    const float acc20 = z0t[kk] + yip0;
                    const float acc21 = z1t[kk] + yip1;
                    float r = dw * (acc21 + xip1);
                    float c = dw * (acc20 + xip0);
                    r = r + 0.5f;
                    c = c + 0.5f;
                    float val = 0.0f;
                    if (!(r < 0.0f) && !(r >= fh) && !(c < 0.0f) && !(c >= fw))
                        val = img[(size_t)(int)r * a.w + (int)c];
                    acc[kk] = acc[kk] + dw2 * val;
                }
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (kk < kn)
            a.vol[((size_t)(a.nz - 1 - (k0 + kk)) * a.ny + j) * a.nx + (a.nx - 1 - i)] = acc[kk];
}

template <int KT>
__global__ __launch_bounds__(256) void fdkd_tilt_kernel(FdkdArgs a)
{
    int bx, by, bz;
    if (!fdk_tile(a.order, a.ntx, a.nty, a.ntz, bx, by, bz)) return;
    int i, j;
    fdk_voxel(a.wx, a.bwx, bx, by, i, j);
    const int k0 = bz * KT;
    if (i >= a.nx || j >= a.ny) return;
    const int kn = min(KT, a.nz - k0);
    float acc[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0f;
    const int2 m = a.mask[j];
    if (i > m.x && i <= m.y) {
        const double x = (double)(a.ox + (float)i * a.sx);
        const double y = (double)(a.oy + (float)j * a.sy);
        const double fw = (double)a.w, fh = (double)a.h;
        for (int p = 0; p < a.np; ++p) {
            const double *t = a.ptab + 8 * p;
            const double xip0 = x * t[0], xip1 = x * t[1], xip2 = x * t[2];
            const double yip0 = y * t[3], yip1 = y * t[4], yip2 = y * t[5];
            const double acc22 = t[7] + yip2;
            const double dw = 1.0 / (acc22 + xip2);
            const double dw2 = dw * dw;
            const float *img = a.img + (size_t)p * a.h * a.w;
            const double *z0t = a.zt0 + (size_t)p * a.nz + k0;
            const double *z1t = a.zt1 + (size_t)p * a.nz + k0;
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) {
                if (kk < kn) {
                    const double acc30 = (z0t[kk] + yip0) + xip0;
                    const double acc31 = (z1t[kk] + yip1) + xip1;
                    double r = acc31 * dw, c = acc30 * dw;
                    r = r + 0.5;
                    c = c + 0.5;
                    float val = 0.0f;
                    if (!(r < 0.0) && !(r >= fh) && !(c < 0.0) && !(c >= fw))
                        val = img[(size_t)(int)r * a.w + (int)c];
                    acc[kk] = (float)((double)acc[kk] + dw2 * (double)val);
                }
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (kk < kn)
            a.vol[((size_t)(a.nz - 1 - (k0 + kk)) * a.ny + j) * a.nx + (a.nx - 1 - i)] = acc[kk];
}

// Thread layout of the FDK kernels: a wave covers 4 x 16 voxel columns (x by
// y) and a block 2 x 2 waves.  Measured at 1024^3 x 720 (fdk_kernel<16,
// false>), columns per wave along x: 64 -> 215 ms, 16 -> 217, 8 -> 163-168,
// 4 -> 142-145, 2 -> 160, 1 -> 229; the mapping only changes which thread
// owns which column (bit-identical volumes).
constexpr int kFdkWaveX = 4, kFdkBlockWavesX = 2;

}  // namespace mr

struct mr_fdk_plan {
    mr_fdk_params p{};
    mr::FdkGeometry g{};
    int device = 0;
    size_t nx = 0, nz = 0, P = 0;
    size_t cbct[4] = {0, 0, 0, 0};
    float origin[3] = {0, 0, 0};
    float ic[2] = {0, 0};
    double icd[2] = {0, 0};   // FDKbp: ic from its own (wrapping) window
    double *d_ptabd = nullptr;
    double *d_zt1d = nullptr;
    int2 *d_mask = nullptr;
    float *d_hs = nullptr;
    float2 *d_tw = nullptr;
    float *d_ptab = nullptr;
    float *d_zt1 = nullptr;
    float *d_w = nullptr;
    float *d_zt0 = nullptr;     // tilt correction: per-slice column terms
    double *d_zt0d = nullptr;
    size_t cbd3 = 0;            // FDKbp's own window row 3 (its tilt position origin)
    size_t cap_np = 0;
};

extern "C" int mr_fdk_window(const mr_fdk_params *params, size_t *cbct_roi)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(params != nullptr && cbct_roi != nullptr, "NULL argument");
    const mr_fdk_params &q = *params;
    MR_REQUIRE(q.roi[2] > q.roi[0], "ROI x1<=x0");
    mr::FdkGeometry g{};
    g.sod = q.sod; g.sdd = q.sdd; g.pixel[0] = q.pixel[0]; g.pixel[1] = q.pixel[1];
    for (int i = 0; i < 3; ++i) g.voxel[i] = q.voxel[i];
    g.ppoint[0] = q.ppoint[0]; g.ppoint[1] = q.ppoint[1]; g.center = q.center;
    g.direction = q.direction;
    for (int i = 0; i < 4; ++i) g.roi[i] = q.roi[i];
    g.projection_roi3 = q.projection_roi3;
    mr::fdk_cbct_roi(g, q.roi[2] - q.roi[0], cbct_roi);
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fdk_plan_create(mr_fdk_plan **plan, const mr_fdk_params *params)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && params != nullptr, "NULL argument");
    *plan = nullptr;
    const mr_fdk_params &q = *params;
    MR_REQUIRE(q.roi[2] > q.roi[0] && q.roi[3] > q.roi[1], "ROI x1<=x0 or y1<=y0");
    MR_REQUIRE(q.sdd > 0 && q.sod > 0 && q.pixel[0] > 0 && q.pixel[1] > 0, "invalid distances/pixel size");
    MR_REQUIRE(q.direction == 0 || q.direction == 1, "direction must be 0 (CW) or 1 (CCW)");
    MR_REQUIRE(q.module == MR_FDK_SINGLE || q.module == MR_FDK_DOUBLE, "module must be MR_FDK_SINGLE or MR_FDK_DOUBLE");
    MR_REQUIRE(!(q.fused && q.module == MR_FDK_DOUBLE), "fused mode is defined for FDKbp_single only");
    MR_REQUIRE(q.filter == 0 || q.filter == 1, "filter must be 0 (reference ramp) or 1 (Shepp-Logan)");
    MR_REQUIRE(q.cosine_weight == 0 || q.cosine_weight == 1, "cosine_weight must be 0 or 1");
    MR_REQUIRE(q.correct_tilt == 0 || q.correct_tilt == 1, "correct_tilt must be 0 or 1");
    MR_REQUIRE(!(q.correct_tilt && q.fused), "fused mode has no tilt correction");
    if (q.correct_tilt)
        MR_REQUIRE(std::isfinite(q.tilt_angle) && std::fabs(q.tilt_angle) < 90.0f && std::isfinite(q.tilt_pivot),
                   "tilt angle must be finite and below 90 degrees, the pivot finite");
    auto *p = new mr_fdk_plan();
    p->p = q;
    mr::FdkGeometry &g = p->g;
    g.sod = q.sod; g.sdd = q.sdd; g.pixel[0] = q.pixel[0]; g.pixel[1] = q.pixel[1];
    for (int i = 0; i < 3; ++i) g.voxel[i] = q.voxel[i];
    g.ppoint[0] = q.ppoint[0]; g.ppoint[1] = q.ppoint[1]; g.center = q.center;
    g.direction = q.direction;
    for (int i = 0; i < 4; ++i) g.roi[i] = q.roi[i];
    g.projection_roi3 = q.projection_roi3;
    if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
    p->nx = q.roi[2] - q.roi[0];
    p->nz = q.roi[3] - q.roi[1];
    p->P = mr::fdk_padwidth(p->nx);
    mr::fdk_cbct_roi(g, p->nx, p->cbct);
    const float U = (float)(q.roi[2] - q.roi[0]);
    const float V = (float)(q.roi[3] - q.roi[1]);
    p->origin[0] = -(U - q.center) * q.voxel[0] - q.voxel[0] / 2;
    p->origin[1] = -(U - q.center) * q.voxel[1] - q.voxel[1] / 2;
    p->origin[2] = -(V - (q.ppoint[1] - (float)q.roi[1])) * q.voxel[2] - q.voxel[2] / 2;
    p->ic[0] = q.ppoint[0] - (float)p->cbct[0];
    p->ic[1] = q.ppoint[1] - (float)p->cbct[1];
    {
        size_t cbd[4];
        mr::fdk_cbct_roi(g, p->nx, cbd, true);
        p->icd[0] = (double)(q.ppoint[0] - (float)cbd[0]);
        p->icd[1] = (double)(q.ppoint[1] - (float)cbd[1]);
        p->cbd3 = cbd[3];
    }
    std::vector<size_t> first, second;
    mr::circle_mask(p->nx, p->nx, q.use_roi != 0, q.matrix_roi, true, first, second,
                    q.correct_tilt ? mr::tilt_mask_shrink(q.tilt_angle, q.tilt_rows) : 0.0f);
    std::vector<int2> m(p->nx);
    for (size_t y = 0; y < p->nx; ++y) { m[y].x = (int)first[y]; m[y].y = (int)second[y]; }
    // ramp_filter_tuned's spectral ramp with its roll-off (fdkbp_single.cpp:906-919)
    const size_t P = p->P;
    std::vector<float> ramp(P);
    const float DEGTORADF = 3.1415926535897932384626f / 180.0f;
    if (q.filter == 1) {
        // Shepp-Logan: |k| / (P/2) * sin(pi k / P) / (pi k / P), evaluated in double
        const double pi = 3.14159265358979323846264338327950288;
        for (size_t i = 0; i < P; ++i) {
            const double k = i < P / 2 ? (double)i : (double)(P - i);
            const double x = pi * k / (double)P;
            ramp[i] = i == 0 ? 0.0f : (float)(k / ((double)P * 0.5) * std::sin(x) / x);
        }
    } else if (q.module == MR_FDK_SINGLE) {
        for (size_t i = 0; i < P; ++i) {
            ramp[i] = i < P / 2 ? (float)(unsigned)i : (float)(unsigned)(P - i);
            ramp[i] *= (0.54 + 0.46 * (std::cos((float)(unsigned)i * DEGTORADF * 360 / (float)(unsigned)P) + 1));
            ramp[i] /= (double)(unsigned)P * 0.5;
        }
    } else {
        // FDKbp's ramp in double (fdkbp.cpp:895-901), rounded to float for the
        // float spectrum
        const double DEGTORAD = 3.14159265358979323846264338327950288 / 180.0;
        for (size_t i = 0; i < P; ++i) {
            double r = i < P / 2 ? (double)i : (double)(P - i);
            r *= (0.54 + 0.46 * (std::cos((double)(unsigned)i * DEGTORAD * 360 / (double)(unsigned)P) + 1));
            r /= (double)(unsigned)P * 0.5;
            ramp[i] = (float)r;
        }
    }
    std::vector<float> tw = mr::make_twiddles(P);
    hipError_t e = hipMalloc(&p->d_mask, sizeof(int2) * p->nx);
    if (e == hipSuccess) e = hipMalloc(&p->d_hs, sizeof(float) * P);
    if (e == hipSuccess) e = hipMalloc(&p->d_tw, sizeof(float2) * P);
    if (e == hipSuccess) e = hipMemcpy(p->d_mask, m.data(), sizeof(int2) * p->nx, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_hs, ramp.data(), sizeof(float) * P, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_tw, tw.data(), sizeof(float) * 2 * P, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mr_fdk_plan_destroy(p);
        return mr::hip_fail(e, "mr_fdk_plan_create", __FILE__, __LINE__);
    }
    *plan = p;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fdk_plan_destroy(mr_fdk_plan *plan)
{
    MR_GUARD_BEGIN
    if (!plan) return MR_OK;
    for (void *ptr : {(void *)plan->d_mask, (void *)plan->d_hs, (void *)plan->d_tw,
                      (void *)plan->d_ptab, (void *)plan->d_zt1, (void *)plan->d_w,
                      (void *)plan->d_ptabd, (void *)plan->d_zt1d, (void *)plan->d_zt0,
                      (void *)plan->d_zt0d})
        if (ptr) (void)hipFree(ptr);
    delete plan;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fdk_plan_window(const mr_fdk_plan *plan, size_t *cbct_roi)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && cbct_roi != nullptr, "NULL argument");
    for (int i = 0; i < 4; ++i) cbct_roi[i] = plan->cbct[i];
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fdk_bp(mr_fdk_plan *plan, float *d_proj, size_t n_proj, size_t win_rows,
                         const float *angles, const float *weights, float *d_volume, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(d_volume != nullptr, "volume pointer is NULL");
    MR_REQUIRE(n_proj == 0 || (d_proj && angles && weights), "NULL projection/angle/weight pointer");
    MR_REQUIRE(win_rows >= 11, "FDK window needs at least 11 detector rows (5-row margins)");
    MR_REQUIRE(n_proj <= 65535, "too many projections for one call");
    hipStream_t s = (hipStream_t)stream;
    const size_t nx = plan->nx, nz = plan->nz, w = nx, h = win_rows;
    const bool dbl = plan->p.module == MR_FDK_DOUBLE;
    const bool tilt = plan->p.correct_tilt != 0;
    if (n_proj > plan->cap_np) {
        for (float **ptr : {&plan->d_ptab, &plan->d_zt1, &plan->d_w, &plan->d_zt0})
            if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
        for (double **ptr : {&plan->d_ptabd, &plan->d_zt1d, &plan->d_zt0d})
            if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
        plan->cap_np = 0;
        MR_HIP(hipMalloc(&plan->d_ptab, sizeof(float) * 8 * n_proj));
        MR_HIP(hipMalloc(&plan->d_zt1, sizeof(float) * nz * n_proj));
        MR_HIP(hipMalloc(&plan->d_w, sizeof(float) * n_proj));
        MR_HIP(hipMalloc(&plan->d_ptabd, sizeof(double) * 8 * n_proj));
        MR_HIP(hipMalloc(&plan->d_zt1d, sizeof(double) * nz * n_proj));
        if (tilt) {
            MR_HIP(hipMalloc(&plan->d_zt0, sizeof(float) * nz * n_proj));
            MR_HIP(hipMalloc(&plan->d_zt0d, sizeof(double) * nz * n_proj));
        }
        plan->cap_np = n_proj;
    }
    // per-projection coefficients (fdkbp_single.cpp:420-456; fdkbp.cpp:430-465)
    const size_t npx = std::max<size_t>(n_proj, 1);
    std::vector<float> ptab(dbl ? 0 : 8 * npx), zt1(dbl ? 0 : nz * npx), zt0(!dbl && tilt ? nz * npx : 0);
    std::vector<double> ptabd(dbl ? 8 * npx : 0), zt1d(dbl ? nz * npx : 0), zt0d(dbl && tilt ? nz * npx : 0);
    const mr_fdk_params &q = plan->p;
    const double kDPi = 3.1415926535897932384626433832795028841971693993751;   // kipl dPi
    for (size_t ip = 0; ip < n_proj; ++ip) {
        if (!dbl) {
            const float *ic = plan->ic;
            float pm[12];
            mr::fdk_proj_matrix(plan->g, angles[ip], pm);
            if (pm[2] != 0.0f || pm[10] != 0.0f) {
                mr::set_error("projection matrix depends on z in x/depth (tilted geometry) - unsupported");
                return MR_ERR_UNSUPPORTED;
            }
            float *t = &ptab[8 * ip];
            t[0] = pm[0] + ic[0] * pm[8];
            t[1] = pm[4] + ic[1] * pm[8];
            t[2] = pm[8];
            t[3] = pm[1] + ic[0] * pm[9];
            t[4] = pm[5] + ic[1] * pm[9];
            t[5] = pm[9];
            t[6] = ic[0] * pm[11] + pm[3];
            t[7] = pm[11];
            const float a1 = pm[6] + ic[1] * pm[10];
            const float b1 = ic[1] * pm[11];
            for (size_t k = 0; k < nz; ++k) {
                const float z = plan->origin[2] + (float)(long)k * plan->p.voxel[2];
                zt1[ip * nz + k] = z * a1 + b1 + pm[7];
                if (tilt) {
                    // slice k's centre of rotation (tan in double, held in a
                    // float) as the column offset (fdkbp_single.cpp:443-448)
                    const float pos = (float)plan->cbct[3] - (float)(long)k - q.tilt_pivot;
                    const float cor = (float)(std::tan((double)(-q.tilt_angle) * kDPi / 180.0) * (double)pos +
                                              (double)q.center);
                    const float pm3 = ((cor - (q.ppoint[0] - (float)q.roi[0])) * q.voxel[0]) / q.pixel[0];
                    zt0[ip * nz + k] = z * (pm[2] + ic[0] * pm[10]) + ic[0] * pm[11] + pm3;
                }
            }
        } else {
            const double *ic = plan->icd;
            double pm[12];
            mr::fdkd_proj_matrix(plan->g, angles[ip], pm);
            if (pm[2] != 0.0 || pm[10] != 0.0) {
                mr::set_error("projection matrix depends on z in x/depth (tilted geometry) - unsupported");
                return MR_ERR_UNSUPPORTED;
            }
            double *t = &ptabd[8 * ip];
            t[0] = pm[0] + ic[0] * pm[8];
            t[1] = pm[4] + ic[1] * pm[8];
            t[2] = pm[8];
            t[3] = pm[1] + ic[0] * pm[9];
            t[4] = pm[5] + ic[1] * pm[9];
            t[5] = pm[9];
            // zip x and depth terms do not depend on the slice when pm[2] and
            // pm[10] are zero: z*(+-0) + ic0*pm11 + pm3 and z*(+-0) + pm11
            const double z0 = (double)(plan->origin[2] + (float)(long)0 * plan->p.voxel[2]);
            t[6] = z0 * (pm[2] + ic[0] * pm[10]) + ic[0] * pm[11] + pm[3];
            t[7] = z0 * pm[10] + pm[11];
            for (size_t k = 0; k < nz; ++k) {
                const double z = (double)(plan->origin[2] + (float)(long)k * plan->p.voxel[2]);
                zt1d[ip * nz + k] = z * (pm[6] + ic[1] * pm[10]) + ic[1] * pm[11] + pm[7];
                if (tilt) {
                    // fdkbp.cpp:452-458: the same in double, from FDKbp's own window
                    const double pos = (double)plan->cbd3 - (double)k - (double)q.tilt_pivot;
                    const double cor = std::tan((double)(-q.tilt_angle) * kDPi / 180.0) * pos + (double)q.center;
                    const double pm3 = ((cor - (double)(q.ppoint[0] - (float)q.roi[0])) * (double)q.voxel[0]) /
                                       (double)q.pixel[0];
                    zt0d[ip * nz + k] = z * (pm[2] + ic[0] * pm[10]) + ic[0] * pm[11] + pm3;
                }
            }
        }
    }
    if (n_proj == 0) {
        // nothing back-projected: the reference leaves the zeroed volume
        int rz = mr::fill_zero(d_volume, nx * nx * nz, s);
        if (rz != MR_OK) return rz;
        MR_HIP(hipStreamSynchronize(s));
        return MR_OK;
    }
    if (!dbl) {
        MR_HIP(hipMemcpyAsync(plan->d_ptab, ptab.data(), sizeof(float) * 8 * n_proj, hipMemcpyHostToDevice, s));
        MR_HIP(hipMemcpyAsync(plan->d_zt1, zt1.data(), sizeof(float) * nz * n_proj, hipMemcpyHostToDevice, s));
        if (tilt)
            MR_HIP(hipMemcpyAsync(plan->d_zt0, zt0.data(), sizeof(float) * nz * n_proj, hipMemcpyHostToDevice, s));
    } else {
        MR_HIP(hipMemcpyAsync(plan->d_ptabd, ptabd.data(), sizeof(double) * 8 * n_proj, hipMemcpyHostToDevice, s));
        MR_HIP(hipMemcpyAsync(plan->d_zt1d, zt1d.data(), sizeof(double) * nz * n_proj, hipMemcpyHostToDevice, s));
        if (tilt)
            MR_HIP(hipMemcpyAsync(plan->d_zt0d, zt0d.data(), sizeof(double) * nz * n_proj, hipMemcpyHostToDevice,
                                  s));
    }
    MR_HIP(hipMemcpyAsync(plan->d_w, weights, sizeof(float) * n_proj, hipMemcpyHostToDevice, s));
    // weight, ramp filter, 1/P, (SOD/SDD)^2, then SOD/SDD (FDKbp_single) or
    // SOD/SDD/nProj (FDKbp)  (fdkreconbase.cpp:247, fdkbp_single.cpp:376-389,
    // fdkbp.cpp:335,413-418)
    const float sod = plan->p.sod, sdd = plan->p.sdd;
    mr::RowFilterArgs fa{};
    fa.rows = d_proj; fa.n_rows = (long)(n_proj * h); fa.n = (int)w; fa.mode = 1;
    fa.insert = (int)((plan->P - w) / 2); fa.taper = nullptr; fa.hs = plan->d_hs; fa.tw = plan->d_tw;
    fa.post0 = 1.0f / (float)plan->P; fa.post1 = (sod * sod) / (sdd * sdd);
    fa.post2 = dbl ? sod / sdd / (float)n_proj : sod / sdd;
    fa.pre_w = plan->d_w; fa.rows_per_proj = (int)h;
    if (plan->p.cosine_weight) {
        fa.cosw = 1;
        fa.cw_u = plan->p.pixel[0] / sdd; fa.cw_v = plan->p.pixel[1] / sdd;
        fa.cw_c0 = dbl ? (float)plan->icd[0] : plan->ic[0];
        fa.cw_r0 = dbl ? (float)plan->icd[1] : plan->ic[1];
    }
    int rc = mr::run_row_filter(fa, (int)plan->P, s);
    if (rc != MR_OK) return rc;
    rc = mr::run_fdk_margin_rows(d_proj, (int)w, (int)h, (long)n_proj, s);
    if (rc != MR_OK) return rc;
    constexpr int KT = 16;
    if (!dbl) {
        MR_REQUIRE(w * h * 4 < (size_t)1 << 31, "FDK window larger than 2^29 samples");
        mr::FdkArgs a{};
        a.img = d_proj; a.vol = d_volume; a.ptab = plan->d_ptab; a.zt1 = plan->d_zt1; a.mask = plan->d_mask;
        a.zt0 = plan->d_zt0;
        a.nx = (int)nx; a.ny = (int)nx; a.nz = (int)nz; a.w = (int)w; a.h = (int)h; a.np = (int)n_proj;
        a.ox = plan->origin[0]; a.oy = plan->origin[1]; a.sx = plan->p.voxel[0]; a.sy = plan->p.voxel[1];
        a.wx = mr::kFdkWaveX;
        a.bwx = mr::kFdkBlockWavesX;
        const int bw = a.bwx * a.wx, bh = 256 / bw;
        a.ntx = (int)((nx + bw - 1) / bw); a.nty = (int)((nx + bh - 1) / bh); a.ntz = (int)((nz + KT - 1) / KT);
        // grid order (blocks in x, y, slice order; measured faster than the
        // XCD-striped order for this kernel)
        a.order = 0;
        const dim3 g((unsigned)a.ntx, (unsigned)a.nty, (unsigned)a.ntz);
        mr::ProfScope prof("fdk_bp", s);
        if (tilt) hipLaunchKernelGGL((mr::fdk_tilt_kernel<KT>), g, dim3(256), 0, s, a);
        else if (plan->p.fused) hipLaunchKernelGGL((mr::fdk_kernel<KT, true>), g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((mr::fdk_kernel<KT, false>), g, dim3(256), 0, s, a);
    } else {
        MR_REQUIRE(w * h * 4 < (size_t)1 << 31, "FDK window larger than 2^29 samples");
        mr::FdkdArgs a{};
        a.img = d_proj; a.vol = d_volume; a.ptab = plan->d_ptabd; a.zt1 = plan->d_zt1d; a.mask = plan->d_mask;
        a.zt0 = plan->d_zt0d;
        a.nx = (int)nx; a.ny = (int)nx; a.nz = (int)nz; a.w = (int)w; a.h = (int)h; a.np = (int)n_proj;
        a.ox = plan->origin[0]; a.oy = plan->origin[1]; a.sx = plan->p.voxel[0]; a.sy = plan->p.voxel[1];
        a.wx = mr::kFdkWaveX;
        a.bwx = mr::kFdkBlockWavesX;
        const int bw = a.bwx * a.wx, bh = 256 / bw;
        a.ntx = (int)((nx + bw - 1) / bw); a.nty = (int)((nx + bh - 1) / bh); a.ntz = (int)((nz + KT - 1) / KT);
        const long ntiles = (long)a.ntx * a.nty * a.ntz;
        // XCD-striped order (measured faster for the double kernel, 1024^3: 258 vs 302 ms)
        a.order = ntiles < (1L << 30) ? 1 : 0;
        const dim3 g = a.order ? dim3((unsigned)(8 * ((ntiles + 7) / 8)))
                               : dim3((unsigned)a.ntx, (unsigned)a.nty, (unsigned)a.ntz);
        mr::ProfScope prof("fdkd_bp", s);
        if (tilt) hipLaunchKernelGGL(mr::fdkd_tilt_kernel<KT>, g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(mr::fdkd_kernel<KT>, g, dim3(256), 0, s, a);
    }
    MR_HIP(hipGetLastError());
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
    MR_GUARD_END
}
