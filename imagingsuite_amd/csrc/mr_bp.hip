// Parallel-beam back-projection on gfx950 (MultiProjBPparallel, MultiProjBP,
// NearestNeighborBP).
//
// Data layout in HBM
//   d_proj  [n_proj][slices][width]    the reference's TImage<float,3> stack
//   sino    [n_used][width][vpad]      weighted, transposed copy (slice fastest),
//                                      vpad = slices rounded up to 4, zero padded:
//                                      exactly the reference's projection buffer
//                                      (StdBackProjectorBase.cpp:169-172)
//   volume  [width(y)][width(x)][slices]  MatrixZXY, slice fastest
//
// bp_kernel: one thread owns one (x, y) voxel column and ZT consecutive
// slices held in registers; it walks every used projection in order, so each
// voxel sees the same sequence of float additions as the reference's z-pillar
// loop (MultiProjBPparallel.cpp:133-163) and the result is bit-identical.
// Geometry tables are read with scalar loads; detector columns are read as
// float4 vectors along the slice axis.
#include "mr_bp.h"
#include "mr_common.h"
#include "mr_filter.h"
#include "mr_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace mr {

// sino[k][u][v] = proj[idx[k]][v][u] * w[idx[k]]  (v < slices), 0 for v >= slices.
__global__ __launch_bounds__(256) void bp_prep_kernel(const float *__restrict__ proj,
                                                      float *__restrict__ sino,
                                                      const int *__restrict__ idx,
                                                      const float *__restrict__ wts, int nu, int nv,
                                                      int vpad)
{
    __shared__ float tile[32][33];
    const int k = blockIdx.z;
    const int p = idx[k];
    const float w = wts[p];
    const int u0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const float *src = proj + (size_t)p * nu * nv;
    for (int r = ty; r < 32; r += 8) {
        const int v = v0 + r, u = u0 + tx;
        tile[r][tx] = (v < nv && u < nu) ? src[(size_t)v * nu + u] * w : 0.0f;
    }
    __syncthreads();
    float *dst = sino + (size_t)k * nu * vpad;
    for (int r = ty; r < 32; r += 8) {
        const int u = u0 + r, v = v0 + tx;
        if (u < nu && v < vpad) dst[(size_t)u * vpad + v] = tile[tx][r];
    }
}

// The same weighting and transposition with 64 x 64 tiles and 16-byte
// global loads and stores (rows of a projection and of the sinogram are
// read and written four floats at a time); width % 4 == 0.
__global__ __launch_bounds__(256) void bp_prep4_kernel(const float *__restrict__ proj,
                                                       float *__restrict__ sino,
                                                       const int *__restrict__ idx,
                                                       const float *__restrict__ wts, int nu, int nv,
                                                       int vpad)
{
    __shared__ float tile[64][65];
    const int k = blockIdx.z;
    const int p = idx[k];
    const float w = wts[p];
    const int u0 = blockIdx.x * 64, v0 = blockIdx.y * 64;
    const int c4 = threadIdx.x & 15, rr = threadIdx.x >> 4;   // 16 x 16
    const float *src = proj + (size_t)p * nu * nv;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = rr + 16 * i, v = v0 + r, u = u0 + 4 * c4;
        float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (v < nv && u < nu) x = *reinterpret_cast<const float4 *>(src + (size_t)v * nu + u);
        tile[r][4 * c4 + 0] = x.x * w;
        tile[r][4 * c4 + 1] = x.y * w;
        tile[r][4 * c4 + 2] = x.z * w;
        tile[r][4 * c4 + 3] = x.w * w;
    }
    __syncthreads();
    float *dst = sino + (size_t)k * nu * vpad;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = rr + 16 * i, u = u0 + r, v = v0 + 4 * c4;
        if (u < nu && v < vpad)
            *reinterpret_cast<float4 *>(dst + (size_t)u * vpad + v) =
                make_float4(tile[4 * c4][r], tile[4 * c4 + 1][r], tile[4 * c4 + 2][r], tile[4 * c4 + 3][r]);
    }
}

// Weighting and transposition of the used projections into the sinogram.
int launch_bp_prep(const float *proj, float *sino, const int *idx, const float *wts, int nu, int nv,
                          int vpad, int nk, hipStream_t s)
{
    ProfScope prof("bp_prep", s);
    if (nu % 4 == 0) {
        dim3 g((unsigned)((nu + 63) / 64), (unsigned)((vpad + 63) / 64), (unsigned)nk);
        hipLaunchKernelGGL(bp_prep4_kernel, g, dim3(256), 0, s, proj, sino, idx, wts, nu, nv, vpad);
    } else {
        dim3 g((unsigned)((nu + 31) / 32), (unsigned)((vpad + 31) / 32), (unsigned)nk);
        hipLaunchKernelGGL(bp_prep_kernel, g, dim3(256), 0, s, proj, sino, idx, wts, nu, nv, vpad);
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// Block -> tile mapping of the tiled kernels.  order 0: the 3-D grid itself.
// order 1: a 1-D grid.  Consecutive blocks are dispatched to consecutive XCDs
// (round robin over the 8 XCDs), so XCD c runs blocks c, c+8, c+16, ...;
// those take the c-th contiguous eighth of a tile list ordered by slice
// chunk, then stripes of sw tile columns, then rows inside a stripe.  The
// blocks resident on one XCD at a time then cover a compact patch of the
// plane, and the projection spans they stage overlap in that XCD's L2.
// Blocks past the end of the list return before touching anything.
__device__ __forceinline__ bool tile_of_block(const BpArgs &a, int &bx, int &by, int &bz)
{
    if (a.order == 0) {
        bx = blockIdx.x;
        by = blockIdx.y;
        bz = blockIdx.z;
        return true;
    }
    const int n = a.ntx * a.nty * a.ntz;
    const int per = (n + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    if (t >= n) return false;
    const int plane = a.ntx * a.nty;
    bz = t / plane;
    const int r = t - bz * plane;
    const int s = r / (a.sw * a.nty);
    const int q = r - s * a.sw * a.nty;
    const int w = min(a.sw, a.ntx - a.sw * s);
    by = q / w;
    bx = a.sw * s + (q - by * w);
    return true;
}

template <int ZT, int INTERP>
__global__ __launch_bounds__(256) void bp_kernel(BpArgs a)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z0 = blockIdx.z * ZT;
    if (x >= a.nu || y >= a.nu) return;
    const int zn = min(ZT, a.nv - z0);          // slices stored by this thread
    const int zl = max(0, min(ZT, a.nz4 - z0));  // slices updated (multiple of 4)
    float *out = a.vol + ((size_t)y * a.nu + x) * a.nv + z0;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = (a.accumulate && z < zn) ? out[z] : 0.0f;
    const int2 m = a.mask[y];
    if (x >= m.x && x < m.y && zl > 0) {
        const float yf = (float)(y + 1);
        const float xf = (float)(x + 1);
        const int num2 = a.nu - 2;
        const float *base = a.sino + z0;
        for (int k = 0; k < a.nk; ++k) {
            const float4 t = a.tab[k];
            const float lsu = t.x + t.y * yf;
            const float u = lsu - t.z * xf;
            const int n = (int)u;
            if (n < 0 || n > num2) continue;
            const float *A = base + ((size_t)k * a.nu + n) * a.vpad;
            if (INTERP == 0) {
                const float wb = fabsf(u - (float)n);
                const float wa = 1.0f - wb;
                const float *B = A + a.bstride;
#pragma unroll
                for (int g = 0; g < ZT / 4; ++g) {
                    if (4 * g < zl) {
                        const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                        const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                        acc[4 * g + 0] = acc[4 * g + 0] + (va.x * wa + vb.x * wb);
                        acc[4 * g + 1] = acc[4 * g + 1] + (va.y * wa + vb.y * wb);
                        acc[4 * g + 2] = acc[4 * g + 2] + (va.z * wa + vb.z * wb);
                        acc[4 * g + 3] = acc[4 * g + 3] + (va.w * wa + vb.w * wb);
                    }
                }
            } else {
#pragma unroll
                for (int g = 0; g < ZT / 4; ++g) {
                    if (4 * g < zl) {
                        const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                        acc[4 * g + 0] = acc[4 * g + 0] + va.x;
                        acc[4 * g + 1] = acc[4 * g + 1] + va.y;
                        acc[4 * g + 2] = acc[4 * g + 2] + va.z;
                        acc[4 * g + 3] = acc[4 * g + 3] + va.w;
                    }
                }
            }
        }
    }
    if ((a.nv & 3) == 0 && zn == ZT) {
#pragma unroll
        for (int g = 0; g < ZT / 4; ++g)
            *reinterpret_cast<float4 *>(out + 4 * g) =
                make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
    } else {
#pragma unroll
        for (int z = 0; z < ZT; ++z)
            if (z < zn) out[z] = acc[z];
    }
}

// Axis-tilt branches of the back-projectors (bCorrectTilt), direct gathers.
// VAR 1, MultiProjBPparallel (MultiProjBPparallel.cpp:49-114):
//   u = ((fStartU + fCos y) - centeroffset) - fSin x, n = (int)u, f = u - n;
//   slice group g (4 slices) weights column n+1 by |f_g - g centerinc|, f_g
//   being f with centerinc subtracted g times;
// VAR 2, MultiProjBP (MultiProjBP.cpp:50-96): u = ((fStartU + fCos y) - fSin x)
//   - centeroffset, group g weights column n+1 by |(u - n) - g centerinc|;
// VAR 3, NearestNeighborBP (NNMultiProjBP.cpp:48-81): u as VAR 2, column n.
// The reference reads columns n and n+1 for every group (its weights are not
// a geometric tilt model) and this kernel does the same, in its float order.
// g counts slice groups from the block's first slice, so a thread whose tile
// starts at group g0 first subtracts centerinc g0 times, as the reference's
// loop has by then.
template <int ZT, int VAR>
__global__ __launch_bounds__(256) void bp_tilt_kernel(BpArgs a)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z0 = blockIdx.z * ZT;
    if (x >= a.nu || y >= a.nu) return;
    const int zn = min(ZT, a.nv - z0);
    const int zl = max(0, min(ZT, a.nz4 - z0));
    const int g0 = z0 >> 2;
    float *out = a.vol + ((size_t)y * a.nu + x) * a.nv + z0;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = (a.accumulate && z < zn) ? out[z] : 0.0f;
    const int2 m = a.mask[y];
    if (x >= m.x && x < m.y && zl > 0) {
        const float yf = (float)(y + 1);
        const float xf = (float)(x + 1);
        const float co = a.tilt_co, ci = a.tilt_ci;
        const int num2 = a.nu - 2;
        const float *base = a.sino + z0;
        for (int k = 0; k < a.nk; ++k) {
            const float4 t = a.tab[k];
            float u;
            if (VAR == 1) {
                const float lsu = (t.x + t.y * yf) - co;
                u = lsu - t.z * xf;
            } else {
                const float lsu = t.x + t.y * yf;
                u = (lsu - t.z * xf) - co;
            }
            const int n = (int)u;
            if (n < 0 || n > num2) continue;
            const float *A = base + ((size_t)k * a.nu + n) * a.vpad;
            if (VAR == 3) {
#pragma unroll
                for (int g = 0; g < ZT / 4; ++g) {
                    if (4 * g < zl) {
                        const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                        acc[4 * g + 0] = acc[4 * g + 0] + va.x;
                        acc[4 * g + 1] = acc[4 * g + 1] + va.y;
                        acc[4 * g + 2] = acc[4 * g + 2] + va.z;
                        acc[4 * g + 3] = acc[4 * g + 3] + va.w;
                    }
                }
                continue;
            }
            const float *B = A + a.bstride;
            const float fr = u - (float)n;
            float f = fr;
            if (VAR == 1)
                for (int j = 0; j < g0; ++j) f = f - ci;
#pragma unroll
            for (int g = 0; g < ZT / 4; ++g) {
                if (4 * g < zl) {
                    const float gz = (float)(g0 + g) * ci;
                    const float wb = VAR == 1 ? fabsf(f - gz) : fabsf(fr - gz);
                    const float wa = 1.0f - wb;
                    const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                    const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                    acc[4 * g + 0] = acc[4 * g + 0] + (va.x * wa + vb.x * wb);
                    acc[4 * g + 1] = acc[4 * g + 1] + (va.y * wa + vb.y * wb);
                    acc[4 * g + 2] = acc[4 * g + 2] + (va.z * wa + vb.z * wb);
                    acc[4 * g + 3] = acc[4 * g + 3] + (va.w * wa + vb.w * wb);
                }
                if (VAR == 1) f = f - ci;
            }
        }
    }
    if ((a.nv & 3) == 0 && zn == ZT) {
#pragma unroll
        for (int g = 0; g < ZT / 4; ++g)
            *reinterpret_cast<float4 *>(out + 4 * g) =
                make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
    } else {
#pragma unroll
        for (int z = 0; z < ZT; ++z)
            if (z < zn) out[z] = acc[z];
    }
}

// LDS-staged variant.  A block owns a 16 x 16 voxel tile and ZT slices.  For
// each chunk of KC projections it stages, per projection, the SPAN detector
// columns the tile can reach (the tile diagonal spans at most 15*sqrt(2)+2
// columns, plus one column of rounding margin below) x ZT slices into LDS, then every thread gathers its two columns
// from LDS with 16-byte reads.  Rows of the staged tile are ZT+4 floats apart,
// which spreads 16 distinct columns over distinct LDS banks.
// EXACT = true keeps the reference's float operation order per voxel
// (bit-identical); EXACT = false folds the interpolation into two FMAs.
constexpr int kBpSpan = 26;

// One 16 x 16 x ZT tile (bx, by, bz) of one volume; shared by the plain grid
// kernel and the persistent batched kernel.  All 256 threads of the block
// must call it together (it synchronises internally).
// PIPE = 0: one projection at a time, slice groups beyond the tile's last
// updated slice skipped (low register use, 8 waves per SIMD).
// PIPE = 1: the projection loop unrolled over the chunk and every slice group
// computed (zeros staged beyond 4*(nv/4)), so the LDS reads of several groups
// and projections are in flight together (about 80 registers, 6 waves).
// TILT (with PIPE = 0 and EXACT): the back-projectors' bCorrectTilt branches
// as in bp_tilt_kernel (1 MultiProjBPparallel, 2 MultiProjBP, 3
// NearestNeighborBP): the detector position shifted by centeroffset and the
// interpolation weight recomputed per slice group.
template <int ZT, int INTERP, bool EXACT, int KC, int PIPE, int TILT = 0>
__device__ __forceinline__ void bp_lds_tile(const BpArgs &a, int bx, int by, int bz,
                                            float (&sm)[KC][kBpSpan][ZT + 4], int (&sb0)[KC])
{
    static_assert(TILT == 0 || (PIPE == 0 && EXACT), "the tilt branches run on the reference-order, rolled path");
    static_assert(TILT != 3 || INTERP == 1, "NearestNeighborBP's tilt branch is the nearest-neighbour form");
    constexpr int T = 16;
    constexpr int SPAN = kBpSpan;
    constexpr int RS = ZT + 4;
    constexpr int G = ZT / 4;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = bx * T, y0 = by * T;
    const int x = x0 + tx, y = y0 + ty;
    const int z0 = bz * ZT;
    const bool in_vol = x < a.nu && y < a.nu;
    const int zn = min(ZT, a.nv - z0);
    const int zl = max(0, min(ZT, a.nz4 - z0));
    float *out = a.vol + ((size_t)min(y, a.nu - 1) * a.nu + min(x, a.nu - 1)) * a.nv + z0;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = (in_vol && a.accumulate && z < zn) ? out[z] : 0.0f;
    bool inside = false;
    if (in_vol) {
        const int2 m = a.mask[y];
        inside = x >= m.x && x < m.y;
    }
    const bool any_inside = __syncthreads_or(inside ? 1 : 0) != 0;
    if (any_inside && zl > 0) {
        const float yf = (float)(y + 1), xf = (float)(x + 1);
        const int num2 = a.nu - 2;
        // Staging: element e = threadIdx.x + 256 i of a chunk is (projection
        // kk, column b, slice group q).  Its LDS slot and its byte offset in
        // the chunk (less the column term) do not change between chunks, so
        // they are computed once here; a chunk then costs one multiply-add
        // per element.  Loads go through a buffer resource over the chunk:
        // slices at or beyond 4*(nv/4) get an offset past its end and read as
        // zeros (the accumulation below runs over every group of the tile and
        // leaves them unchanged), and columns past the detector edge, which no
        // voxel reads, stay inside the buffer or read as zeros.  Chunks too
        // large for 32-bit offsets take the general loop.
        constexpr int NE = (KC * SPAN * G + 255) / 256;
        const int vpad4 = 4 * a.vpad;
        const bool off32 = !a.stage64 && (long)KC * (a.nu + SPAN) * vpad4 < 0x7FFFFFF0L;
        int st_kk[NE], st_b[NE], st_lds[NE];
        unsigned st_off[NE];
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = (int)threadIdx.x + 256 * i;
            const int kk = e / (SPAN * G);
            const int r = e - kk * SPAN * G;
            const int b = r / G, q = r - b * G;
            st_kk[i] = e < KC * SPAN * G ? kk : KC;
            st_b[i] = b;
            st_lds[i] = (kk * SPAN + b) * RS + 4 * q;
            // unsigned: only used when the chunk fits 32-bit offsets (off32)
            st_off[i] = z0 + 4 * q < a.nz4
                            ? (unsigned)kk * (unsigned)a.nu * (unsigned)vpad4 + 4u * (unsigned)(z0 + 4 * q)
                            : 0x80000000u;
        }
        float *smf = &sm[0][0][0];
        for (int kb = 0; kb < a.nk; kb += KC) {
            const int kc = min(KC, a.nk - kb);
            if (threadIdx.x < kc) {
                // lowest detector column any voxel of the tile reaches
                const float4 t = a.tab[kb + threadIdx.x];
                const float ya = (float)(y0 + 1), yb = (float)(y0 + T);
                const float xa = (float)(x0 + 1), xb = (float)(x0 + T);
                int n0;
                if constexpr (TILT == 1) {
                    const float la = (t.x + t.y * ya) - a.tilt_co, lb = (t.x + t.y * yb) - a.tilt_co;
                    n0 = min(min((int)(la - t.z * xa), (int)(la - t.z * xb)),
                             min((int)(lb - t.z * xa), (int)(lb - t.z * xb)));
                } else if constexpr (TILT >= 2) {
                    const float la = t.x + t.y * ya, lb = t.x + t.y * yb, co = a.tilt_co;
                    n0 = min(min((int)((la - t.z * xa) - co), (int)((la - t.z * xb) - co)),
                             min((int)((lb - t.z * xa) - co), (int)((lb - t.z * xb) - co)));
                } else {
                    const float la = t.x + t.y * ya, lb = t.x + t.y * yb;
                    n0 = min(min((int)(la - t.z * xa), (int)(la - t.z * xb)),
                             min((int)(lb - t.z * xa), (int)(lb - t.z * xb)));
                }
                // one bin of margin: an interior voxel may round one ulp below a
                // corner and truncate to the bin below
                sb0[threadIdx.x] = max(n0 - 1, 0);
            }
            __syncthreads();
            if (off32) {
                const float *cb = a.sino + (size_t)kb * a.nu * a.vpad;
                const long rem = (long)(a.nk - kb) * a.nu * vpad4;
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)cb, (short)0, (int)min(rem, 0x7FFFFFF0L), 0x00020000);
                // all loads first, then the LDS writes, so the loads are in
                // flight together; elements of absent projections (last
                // chunk) load zeros from an offset past the buffer, and any
                // offset is safe with a buffer resource
                float4 v[NE];
#pragma unroll
                for (int i = 0; i < NE; ++i) {
                    const unsigned col = (unsigned)(sb0[min(st_kk[i], KC - 1)] + st_b[i]);
                    const unsigned off = st_kk[i] < kc ? col * (unsigned)vpad4 + st_off[i] : 0x80000000u;
                    // the whole 16-byte result is reinterpreted: indexing its
                    // elements makes this compiler load one dword only
                    v[i] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
                }
#pragma unroll
                for (int i = 0; i < NE; ++i)
                    if (st_kk[i] < KC) *reinterpret_cast<float4 *>(smf + st_lds[i]) = v[i];
            } else {
                for (int e = threadIdx.x; e < kc * SPAN * G; e += 256) {
                    const int kk = e / (SPAN * G);
                    const int r = e - kk * SPAN * G;
                    const int b = r / G, q = r - b * G;
                    const int col = sb0[kk] + b;
                    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (col < a.nu && z0 + 4 * q < a.nz4)
                        v = *reinterpret_cast<const float4 *>(
                            a.sino + ((size_t)(kb + kk) * a.nu + col) * a.vpad + z0 + 4 * q);
                    *reinterpret_cast<float4 *>(&sm[kk][b][4 * q]) = v;
                }
            }
            __syncthreads();
            if constexpr (PIPE == 0) {
            if (inside) {
                for (int kk = 0; kk < kc; ++kk) {
                    const float4 t = tab_entry(a.tab, kb + kk);
                    float u;
                    if constexpr (TILT == 1) {
                        const float lsu = (t.x + t.y * yf) - a.tilt_co;
                        u = lsu - t.z * xf;
                    } else if constexpr (TILT >= 2) {
                        const float lsu = t.x + t.y * yf;
                        u = (lsu - t.z * xf) - a.tilt_co;
                    } else {
                        const float lsu = t.x + t.y * yf;
                        u = lsu - t.z * xf;
                    }
                    const int n = (int)u;
                    if (n < 0 || n > num2) continue;
                    const int loc = min(max(n - sb0[kk], 0), SPAN - 2);
                    const float *A = &sm[kk][loc][0];
                    if constexpr (TILT == 1 || TILT == 2) {
                        // per slice group g (from the block's first slice): weight
                        // |f_g - g centerinc| (f_g: f less centerinc g times) or
                        // |(u - n) - g centerinc|
                        const float ci = a.tilt_ci;
                        const int g0 = z0 >> 2;
                        const float fr = u - (float)n;
                        float f = fr;
                        if constexpr (TILT == 1) {
                            // the same subtractions in the same order, eight per trip
                            int j = 0;
                            for (; j + 8 <= g0; j += 8) {
#pragma unroll
                                for (int e = 0; e < 8; ++e) f = f - ci;
                            }
                            for (; j < g0; ++j) f = f - ci;
                        }
                        const float *B = A + RS;
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float gz = (float)(g0 + g) * ci;
                            const float wb = TILT == 1 ? fabsf(f - gz) : fabsf(fr - gz);
                            const float wa = 1.0f - wb;
                            if (4 * g < zl) {
                                const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                                const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                                acc[4 * g + 0] = acc[4 * g + 0] + (va.x * wa + vb.x * wb);
                                acc[4 * g + 1] = acc[4 * g + 1] + (va.y * wa + vb.y * wb);
                                acc[4 * g + 2] = acc[4 * g + 2] + (va.z * wa + vb.z * wb);
                                acc[4 * g + 3] = acc[4 * g + 3] + (va.w * wa + vb.w * wb);
                            }
                            if constexpr (TILT == 1) f = f - ci;
                        }
                    } else if (INTERP == 0) {
                        const float wb = fabsf(u - (float)n);
                        const float wa = 1.0f - wb;
                        const float *B = A + RS;
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            if (4 * g < zl) {
                                const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                                const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                                if (EXACT) {
                                    acc[4 * g + 0] = acc[4 * g + 0] + (va.x * wa + vb.x * wb);
                                    acc[4 * g + 1] = acc[4 * g + 1] + (va.y * wa + vb.y * wb);
                                    acc[4 * g + 2] = acc[4 * g + 2] + (va.z * wa + vb.z * wb);
                                    acc[4 * g + 3] = acc[4 * g + 3] + (va.w * wa + vb.w * wb);
                                } else {
                                    acc[4 * g + 0] = __builtin_fmaf(vb.x, wb, __builtin_fmaf(va.x, wa, acc[4 * g + 0]));
                                    acc[4 * g + 1] = __builtin_fmaf(vb.y, wb, __builtin_fmaf(va.y, wa, acc[4 * g + 1]));
                                    acc[4 * g + 2] = __builtin_fmaf(vb.z, wb, __builtin_fmaf(va.z, wa, acc[4 * g + 2]));
                                    acc[4 * g + 3] = __builtin_fmaf(vb.w, wb, __builtin_fmaf(va.w, wa, acc[4 * g + 3]));
                                }
                            }
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            if (4 * g < zl) {
                                const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                                acc[4 * g + 0] = acc[4 * g + 0] + va.x;
                                acc[4 * g + 1] = acc[4 * g + 1] + va.y;
                                acc[4 * g + 2] = acc[4 * g + 2] + va.z;
                                acc[4 * g + 3] = acc[4 * g + 3] + va.w;
                            }
                        }
                    }
                }
            }
            } else {
            if (inside) {
                // the chunk's staging offsets in scalar registers; the
                // projection loop is unrolled so neighbouring projections'
                // LDS reads and FMAs interleave
                int sbr[KC];
#pragma unroll
                for (int q = 0; q < KC; ++q) sbr[q] = __builtin_amdgcn_readfirstlane(sb0[q]);
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    if (kk >= kc) break;
                    const float4 t = tab_entry(a.tab, kb + kk);
                    const float lsu = t.x + t.y * yf;
                    const float u = lsu - t.z * xf;
                    const int n = (int)u;
                    if (n < 0 || n > num2) continue;
                    const int loc = min(max(n - sbr[kk], 0), SPAN - 2);
                    const float *A = &sm[kk][loc][0];
                    if (INTERP == 0) {
                        const float wb = fabsf(u - (float)n);
                        const float wa = 1.0f - wb;
                        const float *B = A + RS;
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                            const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                            if (EXACT) {
                                acc[4 * g + 0] = acc[4 * g + 0] + (va.x * wa + vb.x * wb);
                                acc[4 * g + 1] = acc[4 * g + 1] + (va.y * wa + vb.y * wb);
                                acc[4 * g + 2] = acc[4 * g + 2] + (va.z * wa + vb.z * wb);
                                acc[4 * g + 3] = acc[4 * g + 3] + (va.w * wa + vb.w * wb);
                            } else {
                                acc[4 * g + 0] = __builtin_fmaf(vb.x, wb, __builtin_fmaf(va.x, wa, acc[4 * g + 0]));
                                acc[4 * g + 1] = __builtin_fmaf(vb.y, wb, __builtin_fmaf(va.y, wa, acc[4 * g + 1]));
                                acc[4 * g + 2] = __builtin_fmaf(vb.z, wb, __builtin_fmaf(va.z, wa, acc[4 * g + 2]));
                                acc[4 * g + 3] = __builtin_fmaf(vb.w, wb, __builtin_fmaf(va.w, wa, acc[4 * g + 3]));
                            }
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                            acc[4 * g + 0] = acc[4 * g + 0] + va.x;
                            acc[4 * g + 1] = acc[4 * g + 1] + va.y;
                            acc[4 * g + 2] = acc[4 * g + 2] + va.z;
                            acc[4 * g + 3] = acc[4 * g + 3] + va.w;
                        }
                    }
                }
            }
            }
            __syncthreads();
        }
    }
    if (!in_vol) return;
    if ((a.nv & 3) == 0 && zn == ZT) {
#pragma unroll
        for (int g = 0; g < G; ++g)
            *reinterpret_cast<float4 *>(out + 4 * g) =
                make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
    } else {
#pragma unroll
        for (int z = 0; z < ZT; ++z)
            if (z < zn) out[z] = acc[z];
    }
}

template <int ZT, int INTERP, bool EXACT, int KC, int PIPE>
__global__ __launch_bounds__(256) void bp_lds_kernel(BpArgs a)
{
    __shared__ float sm[KC][kBpSpan][ZT + 4];
    __shared__ int sb0[KC];
    int bx, by, bz;
    if (!tile_of_block(a, bx, by, bz)) return;
    bp_lds_tile<ZT, INTERP, EXACT, KC, PIPE>(a, bx, by, bz, sm, sb0);
}

// The tilt branches on the same tiles (reference order, rolled loop).
template <int ZT, int INTERP, int KC, int TILT>
__global__ __launch_bounds__(256) void bp_lds_tilt_kernel(BpArgs a)
{
    __shared__ float sm[KC][kBpSpan][ZT + 4];
    __shared__ int sb0[KC];
    int bx, by, bz;
    if (!tile_of_block(a, bx, by, bz)) return;
    bp_lds_tile<ZT, INTERP, true, KC, 0, TILT>(a, bx, by, bz, sm, sb0);
}

// Persistent batched back-projection: a grid sized to the device (every
// block resident at once, so block b runs on XCD b % 8 by the round-robin
// dispatch) pulls work items (stack, tile x, tile y, slice chunk) from eight
// partitions, one per XCD.  Each partition is a contiguous, spatially ordered
// run of the item list with about an eighth of the work, so the blocks of
// one XCD stage overlapping projection spans in its L2.  A block whose
// partition is exhausted moves on to the next one; after at most eight
// empty partitions it leaves, so every block reaches the exit.
template <int ZT, int INTERP, bool EXACT, int KC>
__global__ __launch_bounds__(256) void bp_batch_kernel(const BpArgs *__restrict__ stacks,
                                                       const int4 *__restrict__ items,
                                                       const int *__restrict__ part_begin, int *counters)
{
    __shared__ float sm[KC][kBpSpan][ZT + 4];
    __shared__ int sb0[KC];
    __shared__ int item;
    int part = (int)(blockIdx.x & 7);
    for (;;) {
        if (threadIdx.x == 0) {
            int it = -1;
            for (int tries = 0; tries < 8; ++tries) {
                const int k = atomicAdd(&counters[part], 1);
                const int idx = part_begin[part] + k;
                if (idx < part_begin[part + 1]) {
                    it = idx;
                    break;
                }
                part = (part + 1) & 7;
            }
            item = it;
        }
        __syncthreads();
        const int it = item;
        __syncthreads();
        if (it < 0) break;
        const int4 w = items[it];
        const BpArgs a = stacks[w.x];
        bp_lds_tile<ZT, INTERP, EXACT, KC, ((ZT == 32 && KC == 4) || (ZT == 16 && KC == 8)) ? 1 : 0>(
            a, w.y, w.z, w.w, sm, sb0);
        __syncthreads();
    }
}

// Clears the eight partition counters (a kernel, not a memset, so that the
// launch sequence can be captured in a graph).
__global__ void reset_counter_kernel(int *counter)
{
    if (threadIdx.x < 8) counter[threadIdx.x] = 0;
}

template <int ZT, int INTERP>
static int launch_bp(const BpArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.nu + 15) / 16), (unsigned)((a.nu + 15) / 16),
              (unsigned)((a.nv + ZT - 1) / ZT));
    ProfScope prof("bp_parallel", s);
    hipLaunchKernelGGL((bp_kernel<ZT, INTERP>), grid, dim3(256), 0, s, a);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// Adjoint of the parallel back-projection (the project's own operator for
// gradients through FBP; no reference counterpart).  The forward adds, for
// every used projection k and voxel (x, y) inside the mask (1-based x, y),
//   w_k ((1 - f) S_k[n] + f S_k[n + 1]),  u = (fStartU + fCos y) - fSin x,
//   n = (int)u, f = |u - n|, 0 <= n <= nu - 2
// (MultiProjBPparallel.cpp:130-161); so bin b of projection k receives
// (1 - f) g(x, y) from the voxels with n = b and f g(x, y) from those with
// n = b - 1 (nearest neighbour: g(x, y) from n = b).
//
// Work decomposition: the used projections are sorted by stepping class
// (rows when |sin| >= |cos|, else columns) and direction and cut into chunks
// whose contributing voxels, for every block of kAdjNB detector bins, lie
// within kAdjW - 6 voxels on every line.  A block takes kAdjNB bins x one
// chunk of up to kAdjKA projections x ZT slices, one (bin, projection) per
// thread, and walks the lines (rows or columns) of the volume: each line's
// span of candidate voxels is staged into LDS (zeros outside the circle
// mask) while the previous line is consumed, one barrier per line; the
// global loads of a line are issued two lines ahead, into registers.
//
// On one line the detector position u is monotone along the line with slope
// |sin| (or |cos|) >= 1/sqrt 2, so the voxels a bin receives from form one
// run of at most 3 (the interval of u that maps to it is 2 wide; 1 for
// nearest neighbour).  The thread finds the run's first voxel from the affine
// form (one division; for nearest neighbour then a one-step fix-up), and
// accumulates that window of 3: the weights are the forward's float values
// (the hat below for the linear form, the forward's own operations for
// nearest neighbour), zero for window positions outside the run.  Without
// the fix-up the linear form can miss a voxel at a run's end whose weight is
// within the estimate's rounding of 0 (1024^2 x 64 x 900: 0.02 % of the
// outputs change, by at most 1.2e-7; 9.2 against 9.7 ms).  Bin 1 of the linear form also takes u in
// (-1, 0), which (int) truncates to bin 0 (a run of up to 5), and bin nu - 1
// takes only n = nu - 2: bp_adjoint_edge_kernel computes those two bins.  A
// line whose staged span holds no voxel of the mask is skipped.  The volume is
// read once per chunk; slices at or beyond 4*(nv/4) and dropped projections
// get no gradient (the caller zero-fills the output first).  1024^2 x 64 x
// 900 (all projections): 9.6 ms with the hat against 11.5 ms with the
// forward's operations for every weight.
constexpr int kAdjNB = 16, kAdjKA = 16, kAdjW = 80;

// The forward's weight of the voxel at position p (1-based) of line L for
// bin u, with its float operations (MultiProjBPparallel.cpp:130-161): 0 when
// the voxel does not contribute to u.
template <int INTERP>
__device__ __forceinline__ float adj_weight(float su, float c, float s, bool rows, int L, int p, int u, int num2)
{
    const int x = rows ? p : L, y = rows ? L : p;
    const float lsu = su + c * (float)y;
    const float uu = lsu - s * (float)x;
    const int n = (int)uu;
    const float wb = fabsf(uu - (float)n);
    const bool in = n >= 0 && n <= num2;
    const float w_n = INTERP == 1 ? 1.0f : 1.0f - wb;
    return (in && n == u) ? w_n : ((INTERP == 0 && in && n == u - 1) ? wb : 0.0f);
}

template <int INTERP, int ZT>
__global__ __launch_bounds__(256) void bp_adjoint_kernel(const float *__restrict__ g, float *__restrict__ out,
                                                         const float4 *__restrict__ tab,
                                                         const int2 *__restrict__ mask,
                                                         const int2 *__restrict__ cmask,
                                                         const int *__restrict__ idx,
                                                         const float *__restrict__ wts,
                                                         const int4 *__restrict__ chunks,
                                                         const int *__restrict__ order, int n_chunks, int nbt,
                                                         int nzs, int nu, int nv)
{
    constexpr int NB = kAdjNB, KA = kAdjKA, W = kAdjW, RS = ZT + 4, G = ZT / 4;
    __shared__ float seg[2][W][RS];
    __shared__ float4 ep[2 * KA];   // edge rays: (su, c, s, lower u bound)
    __shared__ float eu[2 * KA];    // edge rays: upper u bound
    extern __shared__ int2 spans[]; // line L (1-based) at [L - 1]: staged positions [lo, hi], 1-based
    const int total = n_chunks * nbt * nzs;
    const int per = (total + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int wi = (b & 7) * per + (b >> 3);
    if (wi >= total) return;
    const int bt = wi % nbt;
    const int rest = wi / nbt;
    const int ci = rest % n_chunks;
    const int zs = rest / n_chunks;
    const int4 ch = chunks[ci];                 // first, count, class
    const int tid = (int)threadIdx.x;
    // lanes 16 q .. 16 q + 15 of a wave take 4 projections x 4 neighbouring
    // bins, whose voxels lie within a few staged rows (rows 16 apart share
    // LDS banks)
    const int ai = ((tid >> 2) & 3) | ((tid >> 4) & 12), bi = (tid & 3) | ((tid >> 2) & 12);
    const int u = bt * NB + bi;
    const int num2 = nu - 2;
    // bins with nothing to receive: past the detector, or (nearest
    // neighbour) the last bin, which no voxel's n reaches
    const bool active = ai < ch.y && u < nu && (INTERP == 0 || u <= num2);
    const int k = order[ch.x + min(ai, ch.y - 1)];   // used-projection index
    const float4 t = tab[k];
    const float su = t.x, c = t.y, s = t.z;
    const bool rows = ch.z == 0;
    const int z0 = zs * ZT;
    // the interval of u whose voxels reach bin u (the run estimate; the
    // weights below decide exactly)
    const float fu = (float)u;
    const float ulo = (INTERP == 0 ? u <= 1 : u == 0) ? -1.0f : (INTERP == 0 ? fu - 1.0f : fu);
    const float uhi = fu + 1.0f;
    const float inv = rows ? 1.0f / s : 1.0f / c;
    if (tid < 2 * KA) {
        // edge ray tid: projection tid / 2, the block's first (even) or last valid (odd) bin
        const int eb = min(bt * NB + ((tid & 1) ? NB - 1 : 0), nu - 1);
        const float4 et = tab[order[ch.x + min(tid >> 1, ch.y - 1)]];
        const float efu = (float)eb;
        ep[tid] = make_float4(et.x, et.y, et.z, eb <= 1 ? -1.0f : efu - 1.0f);
        eu[tid] = efu + 1.0f;
    }
    for (int e = tid; e < 2 * W * RS; e += 256) (&seg[0][0][0])[e] = 0.0f;
    __syncthreads();
    // Per line the positions the block's runs can reach (the edge rays'
    // ranges bound every ray's range), cut to the line's run of mask voxels
    // (row y = L: x in mask[y]; column x = L: y in cmask[x]).  Positions
    // outside it hold no gradient, so only [lo, hi] is staged.
    const int ne = 2 * ch.y;
    const int2 *lmask = rows ? mask : cmask;
    for (int L = tid + 1; L <= nu; L += 256) {
        int lo = 0x7fffffff, hi = -0x7fffffff;
        for (int e = 0; e < ne; ++e) {
            const float4 q = ep[e];
            const float ehi = eu[e];
            const float einv = rows ? 1.0f / q.z : 1.0f / q.y;
            float pa, pb;
            if (rows) {
                const float lsu = q.x + q.y * (float)L;
                pa = (lsu - ehi) * einv;
                pb = (lsu - q.w) * einv;
            } else {
                const float bb = q.z * (float)L - q.x;
                pa = (q.w + bb) * einv;
                pb = (ehi + bb) * einv;
            }
            lo = min(lo, (int)floorf(fminf(pa, pb)) - 1);
            hi = max(hi, (int)ceilf(fmaxf(pa, pb)) + 1);
        }
        const int2 m = lmask[L - 1];   // 0-based [m.x, m.y)
        spans[L - 1] = make_int2(max(lo, m.x + 1), min(hi, m.y));
    }
    __syncthreads();
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    auto live = [&](int L) {
        const int2 sp = spans[L - 1];
        return sp.x <= sp.y;
    };
    // Staging of a line's span in two halves: fetch() issues the global
    // loads of line L into registers, put() writes them to seg[buf] one
    // iteration later, so the loads of line L + 2 are in flight while line L
    // is consumed.  Element e = tid + 256 i is (position j = e / G, slice
    // group e % G).
    constexpr int NE = (W * G + 255) / 256;
    float4 pre[NE];
    // When the whole volume lies within 2^31 bytes the loads go through a
    // buffer resource over it with 32-bit byte offsets, whose per-element
    // part (position j of the span, slice group gq) is computed once here:
    // a line then costs one add per element instead of 64-bit address
    // arithmetic.  Elements past the span get an offset past the range
    // (zeros; put() does not store them).
    const size_t vol_bytes = (size_t)nu * nu * nv * sizeof(float);
    const bool off32 = vol_bytes < 0x7FFFFFF0u;
    const unsigned stepb = rows ? (unsigned)nv * 4u : (unsigned)(off32 ? (size_t)nu * nv * 4 : 0);
    unsigned st_off[NE];
#pragma unroll
    for (int i = 0; i < NE; ++i) {
        const int e = tid + 256 * i;
        const int j = e / G, gq = e - j * G;
        st_off[i] = (unsigned)j * stepb + 4u * (unsigned)(z0 + 4 * gq);
    }
    const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)g, (short)0, off32 ? (int)vol_bytes : 0x7FFFFFF0, 0x00020000);
    auto fetch = [&](int L) {
        const int2 sp = spans[L - 1];
        const int n_el = (min(sp.y - sp.x, W - 1) + 1) * G;
        if (off32) {
            // voxel (position p, line L) at ((y - 1) nu + (x - 1)) nv
            const unsigned lineb = rows ? (unsigned)(L - 1) * (unsigned)nu * stepb : (unsigned)(L - 1) * (unsigned)nv * 4u;
            const unsigned pb = lineb + (unsigned)(sp.x - 1) * stepb;
#pragma unroll
            for (int i = 0; i < NE; ++i) {
                const int e = tid + 256 * i;
                const unsigned off = e < n_el ? pb + st_off[i] : 0x80000000u;
                pre[i] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(vrs, off, 0, 0));
            }
            return;
        }
        const size_t base = rows ? (size_t)(L - 1) * nu * nv : (size_t)(L - 1) * nv;
        const size_t step = rows ? (size_t)nv : (size_t)nu * nv;
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int j = e / G, gq = e - j * G;
            pre[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (e < n_el)
                pre[i] = *reinterpret_cast<const float4 *>(g + base + (size_t)(sp.x + j - 1) * step + z0 + 4 * gq);
        }
    };
    auto put = [&](int L, int buf) {
        const int2 sp = spans[L - 1];
        const int n_el = (min(sp.y - sp.x, W - 1) + 1) * G;
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int j = e / G, gq = e - j * G;
            if (e < n_el) *reinterpret_cast<float4 *>(&seg[buf][j][4 * gq]) = pre[i];
        }
    };
    // weight of the voxel at position p of line L for this thread's bin (0
    // when it does not contribute), with the forward's float operations
    auto weight = [&](int L, int p) -> float { return adj_weight<INTERP>(su, c, s, rows, L, p, u, num2); };
    // The same weight for bins 0 and 2 .. nu - 2 of the linear form: the hat
    // 1 - |u' - u| at the forward's position u'.  For those bins it is the
    // forward's float value: u' - u and 1 - |u' - u| are exact (Sterbenz,
    // u' in (u - 1, u + 1) and u >= 2; 1 - |u'| itself for bin 0), and the
    // range tests on n cannot fail.  Bin 1 (which also takes u' in (-1, 0),
    // weight |u'|) and bin nu - 1 (n <= nu - 2 only) differ: this kernel does
    // not write them, bp_adjoint_edge_kernel does.
    auto hat = [&](int L, int p) -> float {
        const int x = rows ? p : L, y = rows ? L : p;
        const float lsu = su + c * (float)y;
        const float uu = lsu - s * (float)x;
        return fmaxf(0.0f, 1.0f - fabsf(uu - fu));
    };
    // adds the voxel at position p of line L (staged in seg[buf]) with weight w
    auto add = [&](int L, int p, int buf, const int2 sp, float w) {
        // staged positions only (a position outside [lo, hi] is outside the
        // mask or past the window's run: weight 0, and a staged value is read)
        int j = p - sp.x;
        const bool out = j < 0 || j > min(sp.y - sp.x, W - 1);
        // a zero-weight position is skipped: its lanes read nothing from LDS
        // (the staged values are finite and the sums never -0, so the result
        // is the same)
        if (out || w == 0.0f) return;
        const float *v = &seg[buf][j][0];
#pragma unroll
        for (int gq = 0; gq < G; ++gq) {
            const float4 x4 = *reinterpret_cast<const float4 *>(v + 4 * gq);
            acc[4 * gq + 0] = __builtin_fmaf(w, x4.x, acc[4 * gq + 0]);
            acc[4 * gq + 1] = __builtin_fmaf(w, x4.y, acc[4 * gq + 1]);
            acc[4 * gq + 2] = __builtin_fmaf(w, x4.z, acc[4 * gq + 2]);
            acc[4 * gq + 3] = __builtin_fmaf(w, x4.w, acc[4 * gq + 3]);
        }
    };
    // the run's window on line L (first voxel: the estimate, one step down
    // if the voxel before it contributes, one step up if it does not)
    auto consume = [&](int L, int buf, auto wf) {
        float pa, pb;
        if (rows) {
            const float lsu = su + c * (float)L;
            pa = (lsu - uhi) * inv;
            pb = (lsu - ulo) * inv;
        } else {
            const float bb = s * (float)L - su;
            pa = (ulo + bb) * inv;
            pb = (uhi + bb) * inv;
        }
        const int est = (int)ceilf(fminf(pa, pb));
        const int2 sp = spans[L - 1];
        if constexpr (INTERP == 1) {
            const int p0 = wf(L, est - 1) != 0.0f ? est - 1 : (wf(L, est) != 0.0f ? est : est + 1);
            // one window position at a time: 112 instead of 146 registers, four
            // waves per SIMD (1024^2 x 64 x 900: 11.5 against 12.1 ms)
#pragma unroll 1
            for (int i = 0; i < 3; ++i) add(L, p0 + i, buf, sp, wf(L, p0 + i));
        } else {
            // the window est .. est + 2 holds the run but for a voxel at one
            // of its ends whose weight is within the estimate's rounding of 0
            // (below 1e-4); the weights are the forward's, so the pair stays
            // an exact transpose up to such terms
#pragma unroll 1
            for (int i = 0; i < 3; ++i) add(L, est + i, buf, sp, wf(L, est + i));
        }
    };
    if (live(1)) {
        fetch(1);
        put(1, 0);
    }
    if (nu >= 2 && live(2)) fetch(2);
    __syncthreads();
    for (int L = 1; L <= nu; ++L) {
        const int buf = (L - 1) & 1;
        // line L + 1: loaded during the previous line, into the other buffer
        if (L < nu && live(L + 1)) put(L + 1, buf ^ 1);
        if (L + 2 <= nu && live(L + 2)) fetch(L + 2);
        if (active && live(L)) {
            if constexpr (INTERP == 1) consume(L, buf, weight);
            else consume(L, buf, hat);
        }
        __syncthreads();
    }
    // bins 1 and nu - 1 of the linear form are bp_adjoint_edge_kernel's
    if (!active || (INTERP == 0 && (u == 1 || u == nu - 1))) return;
    const float w = wts[idx[k]];
    float *o = out + (size_t)idx[k] * nv * nu + u;
#pragma unroll
    for (int z = 0; z < ZT; ++z) o[(size_t)(z0 + z) * nu] = w * acc[z];
}

// Bins 1 and nu - 1 of the linear form, whose weights the hat of
// bp_adjoint_kernel does not reproduce.  One wave per (used projection, edge
// bin), a lane per slice: the wave walks the lines in bp_adjoint_kernel's
// order (the projection's stepping class, lines ascending, positions
// ascending within a line, voxels of the mask only) and adds each
// contributing voxel with the forward's weight by one FMA, the same float
// operations in the same order as the other bins get there (the additions of
// zero weights that kernel makes and this one skips leave a sum unchanged).
// The line walk is the same for every lane, so the mask and the weights are
// wave-uniform; 4 lines are gathered before their FMAs so that their loads
// are in flight together.  Writes only bins 1 and nu - 1 and may run
// concurrently with bp_adjoint_kernel.
__global__ __launch_bounds__(64) void bp_adjoint_edge_kernel(const float *__restrict__ g, float *__restrict__ out,
                                                             const float4 *__restrict__ tab,
                                                             const int2 *__restrict__ mask,
                                                             const int2 *__restrict__ cmask,
                                                             const int *__restrict__ idx,
                                                             const float *__restrict__ wts, int nu, int nv, int nz4)
{
    constexpr int kLines = 4, kPos = 8;   // lines per gather; window positions est - 2 .. est + 5
    const int k = (int)blockIdx.x >> 1, e = (int)blockIdx.x & 1;
    const int u = e == 0 ? 1 : nu - 1;
    if (u >= nu || (e == 1 && u == 1)) return;
    const float4 tb = tab[k];
    const float su = tb.x, c = tb.y, s = tb.z;
    const bool rows = fabsf(s) >= fabsf(c);
    const float fu = (float)u;
    const float ulo = u <= 1 ? -1.0f : fu - 1.0f, uhi = fu + 1.0f;
    const float inv = rows ? 1.0f / s : 1.0f / c;
    const int2 *lmask = rows ? mask : cmask;
    const float wk = wts[idx[k]];
    float *o = out + (size_t)idx[k] * nv * nu + u;
    for (int z = (int)threadIdx.x; z - (int)threadIdx.x < nz4; z += 64) {
        const bool zin = z < nz4;
        float acc = 0.0f;
        for (int L0 = 1; L0 <= nu; L0 += kLines) {
            float wv[kLines][kPos], xv[kLines][kPos];
#pragma unroll
            for (int l = 0; l < kLines; ++l) {
                const int L = L0 + l;
                int lo = 1, hi = 0;
                if (L <= nu) {
                    const int2 m = lmask[L - 1];   // positions m.x + 1 .. m.y (1-based)
                    float pa, pb;
                    if (rows) {
                        const float lsu = su + c * (float)L;
                        pa = (lsu - uhi) * inv;
                        pb = (lsu - ulo) * inv;
                    } else {
                        const float bb = s * (float)L - su;
                        pa = (ulo + bb) * inv;
                        pb = (uhi + bb) * inv;
                    }
                    // the run (up to 5 voxels for bin 1, 3 otherwise) lies in est - 1 .. est + 4
                    const int est = (int)ceilf(fminf(pa, pb));
                    lo = max(est - 2, m.x + 1);
                    hi = min(est - 2 + kPos - 1, m.y);
                }
#pragma unroll
                for (int i = 0; i < kPos; ++i) {
                    const int p = lo + i;
                    const bool inr = p <= hi;
                    const float w = inr ? adj_weight<0>(su, c, s, rows, L, p, u, nu - 2) : 0.0f;
                    wv[l][i] = w;
                    xv[l][i] = 0.0f;
                    if (w != 0.0f && zin) {
                        const size_t off = rows ? ((size_t)(L - 1) * nu + (p - 1)) * nv
                                                : ((size_t)(p - 1) * nu + (L - 1)) * nv;
                        xv[l][i] = g[off + z];
                    }
                }
            }
#pragma unroll
            for (int l = 0; l < kLines; ++l)
#pragma unroll
                for (int i = 0; i < kPos; ++i)
                    if (wv[l][i] != 0.0f) acc = __builtin_fmaf(wv[l][i], xv[l][i], acc);
        }
        if (zin) o[(size_t)z * nu] = wk * acc;
    }
}

// Grid of the tiled kernel (16 x 16 voxel tiles): the XCD-striped 1-D grid
// with stripes of 4 tile columns (order 1, see tile_of_block), or the 3-D
// tile grid when the tile count would overflow the int tile numbers.
static dim3 tiled_grid(BpArgs &a, int ZT)
{
    a.ntx = a.nty = (a.nu + 15) / 16;
    a.ntz = (a.nv + ZT - 1) / ZT;
    const long n = (long)a.ntx * a.nty * a.ntz;
    a.order = n < (1L << 30) ? 1 : 0;
    a.sw = 4;
    if (a.order == 0) return dim3((unsigned)a.ntx, (unsigned)a.nty, (unsigned)a.ntz);
    return dim3((unsigned)(8 * ((n + 7) / 8)));
}

template <int ZT, int INTERP, bool EXACT, int KC, int PIPE>
static int launch_bp_lds(BpArgs a, hipStream_t s)
{
    const dim3 grid = tiled_grid(a, ZT);
    ProfScope prof("bp_parallel", s);
    hipLaunchKernelGGL((bp_lds_kernel<ZT, INTERP, EXACT, KC, PIPE>), grid, dim3(256), 0, s, a);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// Tilt branches on the LDS-staged tiles (the reference order, rolled loop).
template <int ZT, int KC>
static int launch_bp_lds_tilt(BpArgs a, hipStream_t s)
{
    const dim3 grid = tiled_grid(a, ZT);
    ProfScope prof("bp_parallel_tilt", s);
    if (a.tilt == 1) hipLaunchKernelGGL((bp_lds_tilt_kernel<ZT, 0, KC, 1>), grid, dim3(256), 0, s, a);
    else if (a.tilt == 2) hipLaunchKernelGGL((bp_lds_tilt_kernel<ZT, 0, KC, 2>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((bp_lds_tilt_kernel<ZT, 1, KC, 3>), grid, dim3(256), 0, s, a);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

template <int ZT, int KC, int PIPE>
static int launch_bp_lds_mode(const BpArgs &a, int interp, int exact, hipStream_t s)
{
    if (interp == 1) return launch_bp_lds<ZT, 1, true, KC, PIPE>(a, s);
    return exact ? launch_bp_lds<ZT, 0, true, KC, PIPE>(a, s) : launch_bp_lds<ZT, 0, false, KC, PIPE>(a, s);
}

// Slice tile and projection chunk of the LDS kernel: 32 slices with chunks
// of 4 projections (the projection loop unrolled over the chunk; measured
// 104-107 ms at 2048^2 x 256 x 1500 against 146 ms for 16 x 16 x 16 tiles
// with 16-projection chunks); 16 slices with chunks of 8 when the stack has
// fewer than 32 slices or the grid would have fewer than 1024 blocks (4 per
// CU; 256^2 x 64: 0.090 instead of 0.114 ms); 8 slices with chunks of 16
// (rolled loop) below 16 slices.
static void lds_shape(int nu, int nv, int &zt, int &kc, int &pipe)
{
    zt = nv >= 32 ? 32 : (nv >= 16 ? 16 : 8);
    const long tiles = (long)((nu + 15) / 16) * ((nu + 15) / 16);
    if (zt == 32 && tiles * ((nv + 31) / 32) < 1024) zt = 16;
    kc = zt == 32 ? 4 : (zt == 16 ? 8 : 16);
    pipe = zt == 8 ? 0 : 1;
}

// impl 0: direct global gathers (MultiProjBP's second column read at its own
// offset, which differs from the staged layout when slices % 4 != 0);
// impl 1: LDS-staged tiles.  exact: the reference's float operation order
// (otherwise fused multiply-adds).
template <int ZT>
static int launch_bp_tilt(const BpArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.nu + 15) / 16), (unsigned)((a.nu + 15) / 16), (unsigned)((a.nv + ZT - 1) / ZT));
    ProfScope prof("bp_parallel_tilt", s);
    if (a.tilt == 1) hipLaunchKernelGGL((bp_tilt_kernel<ZT, 1>), grid, dim3(256), 0, s, a);
    else if (a.tilt == 2) hipLaunchKernelGGL((bp_tilt_kernel<ZT, 2>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((bp_tilt_kernel<ZT, 3>), grid, dim3(256), 0, s, a);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int run_bp(const BpArgs &a, int interp, int impl, int exact, hipStream_t s)
{
    if (a.tilt != 0) {
        MR_REQUIRE(a.tilt >= 1 && a.tilt <= 3, "unknown tilt branch");
        if (impl == 1) {
            int zt, kc, pipe;
            lds_shape(a.nu, a.nv, zt, kc, pipe);
            if (zt == 32) return launch_bp_lds_tilt<32, 4>(a, s);
            if (zt == 16) return launch_bp_lds_tilt<16, 8>(a, s);
            return launch_bp_lds_tilt<8, 16>(a, s);
        }
        if (a.nv >= 16) return launch_bp_tilt<16>(a, s);
        if (a.nv >= 8) return launch_bp_tilt<8>(a, s);
        return launch_bp_tilt<4>(a, s);
    }
    if (impl == 0) {
        if (a.nv >= 16) return interp == 0 ? launch_bp<16, 0>(a, s) : launch_bp<16, 1>(a, s);
        if (a.nv >= 8) return interp == 0 ? launch_bp<8, 0>(a, s) : launch_bp<8, 1>(a, s);
        return interp == 0 ? launch_bp<4, 0>(a, s) : launch_bp<4, 1>(a, s);
    }
    int zt, kc, pipe;
    lds_shape(a.nu, a.nv, zt, kc, pipe);
    if (zt == 32) return launch_bp_lds_mode<32, 4, 1>(a, interp, exact, s);
    if (zt == 16) return launch_bp_lds_mode<16, 8, 1>(a, interp, exact, s);
    return launch_bp_lds_mode<8, 16, 0>(a, interp, exact, s);
}

// The kernel run_bp launches for these settings, e.g.
// "bp_lds_kernel<32,0,false,4,1> order=1 sw=4".
std::string describe_bp(int nu, int nv, int interp, int impl, int exact, int tilt)
{
    if (tilt != 0 && impl == 1) {
        int zt, kc, pipe;
        lds_shape(nu, nv, zt, kc, pipe);
        const long n = (long)((nu + 15) / 16) * ((nu + 15) / 16) * ((nv + zt - 1) / zt);
        return "bp_lds_tilt_kernel<" + std::to_string(zt) + "," + (tilt == 3 ? "1" : "0") + "," + std::to_string(kc) +
               "," + std::to_string(tilt) + ">" + (n < (1L << 30) ? " order=1 sw=4" : " order=0");
    }
    if (tilt != 0) {
        const int zt = nv >= 16 ? 16 : (nv >= 8 ? 8 : 4);
        return "bp_tilt_kernel<" + std::to_string(zt) + "," + std::to_string(tilt) + ">";
    }
    if (impl == 0) {
        const int zt = nv >= 16 ? 16 : (nv >= 8 ? 8 : 4);
        return "bp_kernel<" + std::to_string(zt) + "," + std::to_string(interp) + ">";
    }
    int zt, kc, pipe;
    lds_shape(nu, nv, zt, kc, pipe);
    const long n = (long)((nu + 15) / 16) * ((nu + 15) / 16) * ((nv + zt - 1) / zt);
    return "bp_lds_kernel<" + std::to_string(zt) + "," + std::to_string(interp) + "," +
           ((interp == 1 || exact) ? "true" : "false") + "," + std::to_string(kc) + "," + std::to_string(pipe) +
           ">" + (n < (1L << 30) ? " order=1 sw=4" : " order=0");
}

}  // namespace mr

struct mr_bp_plan {
    mr_bp_params p{};
    int device = 0;
    size_t vpad = 0;
    std::vector<size_t> first, second;
    int2 *d_mask = nullptr;   // [first, second) per row y
    int2 *d_cmask = nullptr;  // [y0, y1) per column x (the adjoint's; made on first use)
    float *d_sino = nullptr;
    size_t sino_cap = 0;   // floats
    // d_tab | d_idx | d_w live in one allocation, laid out as the pinned
    // staging buffer, so one copy uploads all three (table_layout)
    void *d_tables = nullptr;
    size_t tables_cap = 0; // bytes
    float4 *d_tab = nullptr;
    int *d_idx = nullptr;
    float *d_w = nullptr;
    int impl = 1;  // 1 LDS-staged kernel, 0 direct-gather kernel
    int stage64 = 0;  // 1: 64-bit staging addresses (MR_BP_STAGE64 debug override)
    size_t loaded_n = 0, loaded_nk = 0;   // projections / used projections of the resident tables
    bool loaded = false;
    void *d_adj = nullptr;                // adjoint chunks (int4) and projection order (int)
    size_t adj_cap = 0;                   // bytes
    void *h_stage = nullptr;              // pinned host staging of the tables: tab | idx | w
    size_t h_stage_cap = 0;               // bytes
    hipEvent_t stage_ev = nullptr;        // recorded after the last upload from h_stage
    hipEvent_t done_ev = nullptr;         // recorded after the last mr_bp_parallel's kernels
    hipStream_t side = nullptr;           // the adjoint's edge-bin kernel (default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int tilt = 0;                         // BpArgs::tilt branch (0: no tilt correction)
    size_t tilt_row0 = 0;                 // first detector row of the next block
};

namespace {

// Byte offsets of the used-projection indices and of the weights after the
// nt table entries (float4), and the total size for n_proj weights.
struct TableLayout {
    size_t idx, w, bytes;
};
inline TableLayout table_layout(size_t nt, size_t n_proj)
{
    TableLayout l;
    l.idx = sizeof(float4) * nt;
    l.w = l.idx + sizeof(int) * ((nt + 3) & ~size_t(3));
    l.bytes = l.w + sizeof(float) * n_proj;
    return l;
}

int ensure_capacity(mr_bp_plan *p, size_t n_proj, size_t n_used)
{
    const size_t need = std::max<size_t>(n_used, 1) * p->p.width * p->vpad;
    const size_t nt = std::max<size_t>(n_used, 1);
    const TableLayout tl = table_layout(nt, std::max<size_t>(n_proj, 1));
    // a buffer about to be replaced may still be read by the last call's kernels
    if (p->done_ev && (need > p->sino_cap || tl.bytes > p->tables_cap))
        MR_HIP(hipEventSynchronize(p->done_ev));
    if (need > p->sino_cap) {
        if (p->d_sino) (void)hipFree(p->d_sino);
        p->d_sino = nullptr;
        p->sino_cap = 0;
        MR_HIP(hipMalloc(&p->d_sino, need * sizeof(float)));
        p->sino_cap = need;
    }
    if (tl.bytes > p->tables_cap) {
        if (p->d_tables) (void)hipFree(p->d_tables);
        p->d_tables = nullptr;
        p->tables_cap = 0;
        MR_HIP(hipMalloc(&p->d_tables, tl.bytes));
        p->tables_cap = tl.bytes;
    }
    char *b = static_cast<char *>(p->d_tables);
    p->d_tab = reinterpret_cast<float4 *>(b);
    p->d_idx = reinterpret_cast<int *>(b + tl.idx);
    p->d_w = reinterpret_cast<float *>(b + tl.w);
    return MR_OK;
}

}  // namespace

extern "C" int mr_bp_plan_create(mr_bp_plan **plan, const mr_bp_params *params)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && params != nullptr, "NULL argument");
    *plan = nullptr;
    const mr_bp_params &q = *params;
    MR_REQUIRE(q.width >= 2 && q.width <= 65536, "width must be in [2, 65536]");
    MR_REQUIRE(q.slices >= 1 && q.slices <= 65536, "slices must be in [1, 65536]");
    MR_REQUIRE(q.interp == MR_INTERP_LINEAR || q.interp == MR_INTERP_NEAREST, "unknown interpolation");
    MR_REQUIRE(q.direction == 0 || q.direction == 1, "direction must be 0 (CW) or 1 (CCW)");
    MR_REQUIRE(q.buffer_size >= 0, "buffer_size must be >= 0");
    if (q.use_roi) MR_REQUIRE(q.roi[0] <= q.roi[2] && q.roi[1] <= q.roi[3], "matrix ROI x1<x0 or y1<y0");
    MR_REQUIRE(q.correct_tilt == 0 || q.correct_tilt == 1, "correct_tilt must be 0 or 1");
    if (q.correct_tilt)
        MR_REQUIRE(std::isfinite(q.tilt_angle) && std::fabs(q.tilt_angle) < 90.0f && std::isfinite(q.tilt_pivot),
                   "tilt angle must be finite and below 90 degrees, the pivot finite");
    auto *p = new mr_bp_plan();
    p->p = q;
    if (q.correct_tilt) {
        p->tilt = q.interp == MR_INTERP_NEAREST ? 3 : (q.multiprojbp ? 2 : 1);
        p->tilt_row0 = q.tilt_row0;
    }
    if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
    p->vpad = (q.slices + 3) & ~(size_t)3;
    // MultiProjBP's B-column offset only differs from the padded column when
    // slices % 4 != 0; that case stays on the direct-gather kernel.
    p->impl = (q.multiprojbp && (q.slices & 3)) ? 0 : 1;
    // debug override, read once per plan: MR_BP_STAGE64=1 makes the LDS
    // kernel stage through 64-bit addresses, the path that chunks too large
    // for 32-bit buffer offsets take (tested bit-identical to the default)
    if (const char *e = std::getenv("MR_BP_STAGE64")) p->stage64 = std::atoi(e) == 1 ? 1 : 0;
    mr::circle_mask(q.width, q.width, q.use_roi != 0, q.roi, false, p->first, p->second,
                    q.correct_tilt ? mr::tilt_mask_shrink(q.tilt_angle, q.tilt_rows) : 0.0f);
    std::vector<int2> m(q.width);
    for (size_t y = 0; y < q.width; ++y) {
        // reference loop x = first+1 .. second (1-based) == [first, second) 0-based
        m[y].x = (int)p->first[y];
        m[y].y = (int)std::max(p->first[y], p->second[y]);
    }
    hipError_t e = hipMalloc(&p->d_mask, sizeof(int2) * q.width);
    if (e == hipSuccess) e = hipMemcpy(p->d_mask, m.data(), sizeof(int2) * q.width, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mr_bp_plan_destroy(p);
        return mr::hip_fail(e, "mr_bp_plan_create", __FILE__, __LINE__);
    }
    *plan = p;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_bp_plan_destroy(mr_bp_plan *plan)
{
    MR_GUARD_BEGIN
    if (!plan) return MR_OK;
    if (plan->done_ev) (void)hipEventSynchronize(plan->done_ev);
    if (plan->d_mask) (void)hipFree(plan->d_mask);
    if (plan->d_cmask) (void)hipFree(plan->d_cmask);
    if (plan->d_sino) (void)hipFree(plan->d_sino);
    if (plan->d_tables) (void)hipFree(plan->d_tables);
    if (plan->d_adj) (void)hipFree(plan->d_adj);
    if (plan->stage_ev) {
        (void)hipEventSynchronize(plan->stage_ev);
        (void)hipEventDestroy(plan->stage_ev);
    }
    if (plan->done_ev) (void)hipEventDestroy(plan->done_ev);
    if (plan->side) (void)hipStreamSynchronize(plan->side);
    if (plan->ev_fork) (void)hipEventDestroy(plan->ev_fork);
    if (plan->ev_join) (void)hipEventDestroy(plan->ev_join);
    if (plan->side) (void)hipStreamDestroy(plan->side);
    if (plan->h_stage) (void)hipHostFree(plan->h_stage);
    delete plan;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_bp_plan_kernel(const mr_bp_plan *plan, char *name, size_t len)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && name != nullptr && len > 0, "NULL argument");
    const std::string d = mr::describe_bp((int)plan->p.width, (int)plan->p.slices, plan->p.interp, plan->impl,
                                          plan->p.fused ? 0 : 1, plan->tilt);
    std::snprintf(name, len, "%s", d.c_str());
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_bp_plan_mask(const mr_bp_plan *plan, size_t *first, size_t *second)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && first != nullptr && second != nullptr, "NULL argument");
    std::copy(plan->first.begin(), plan->first.end(), first);
    std::copy(plan->second.begin(), plan->second.end(), second);
    return MR_OK;
    MR_GUARD_END
}

namespace mr {

int bp_load_tables(mr_bp_plan *plan, size_t n_proj, const float *angles, const float *weights, hipStream_t s)
{
    const mr_bp_params &q = plan->p;
    MR_REQUIRE(n_proj == 0 || (angles && weights), "NULL angle/weight pointer");
    std::vector<size_t> used = mr::used_projections(n_proj, (size_t)q.buffer_size);
    const size_t nk = used.size();
    MR_REQUIRE(nk <= 65535, "at most 65535 projections per call");
    plan->loaded = false;
    int rc = ensure_capacity(plan, n_proj, nk);
    if (rc != MR_OK) return rc;
    // the tables and the staged sinogram are overwritten below: on a stream
    // other than the last call's, wait for that call's kernels first
    if (plan->done_ev) MR_HIP(hipStreamWaitEvent(s, plan->done_ev, 0));
    if (nk > 0) {
        std::vector<float> tab;
        mr::bp_tables(angles, used.data(), nk, q.direction, q.center, q.width / 2, tab);
        // The tables go through the plan's pinned staging buffer, so the
        // copies are asynchronous and the stream is not drained here (the
        // filter before and the back-projection after stay back to back);
        // the previous upload from that buffer must have finished reading it.
        // Staging and device buffer share one layout (table_layout), so a
        // single copy uploads the table, the indices and the weights.
        const TableLayout tl = table_layout(nk, n_proj);
        if (plan->stage_ev) MR_HIP(hipEventSynchronize(plan->stage_ev));
        else MR_HIP(hipEventCreateWithFlags(&plan->stage_ev, hipEventDisableTiming));
        if (tl.bytes > plan->h_stage_cap) {
            if (plan->h_stage) (void)hipHostFree(plan->h_stage);
            plan->h_stage = nullptr;
            plan->h_stage_cap = 0;
            MR_HIP(hipHostMalloc(&plan->h_stage, tl.bytes, hipHostMallocDefault));
            plan->h_stage_cap = tl.bytes;
        }
        char *h = static_cast<char *>(plan->h_stage);
        std::memcpy(h, tab.data(), sizeof(float) * tab.size());
        int *hi = reinterpret_cast<int *>(h + tl.idx);
        for (size_t k = 0; k < nk; ++k) hi[k] = (int)used[k];
        for (size_t k = nk; k < (tl.w - tl.idx) / sizeof(int); ++k) hi[k] = 0;
        std::memcpy(h + tl.w, weights, sizeof(float) * n_proj);
        MR_HIP(hipMemcpyAsync(plan->d_tables, h, tl.bytes, hipMemcpyHostToDevice, s));
        MR_HIP(hipEventRecord(plan->stage_ev, s));
    }
    plan->loaded_n = n_proj;
    plan->loaded_nk = nk;
    plan->loaded = true;
    return MR_OK;
}

int bp_enqueue(mr_bp_plan *plan, const float *d_proj, float *d_volume, int accumulate, hipStream_t s)
{
    MR_REQUIRE(plan->loaded, "no projection tables loaded");
    const mr_bp_params &q = plan->p;
    const size_t nk = plan->loaded_nk;
    MR_REQUIRE(nk == 0 || d_proj != nullptr, "NULL projection pointer");
    if (nk > 0) {
        const int rc = mr::launch_bp_prep(d_proj, plan->d_sino, plan->d_idx, plan->d_w, (int)q.width,
                                          (int)q.slices, (int)plan->vpad, (int)nk, s);
        if (rc != MR_OK) return rc;
    }
    mr::BpArgs a{};
    a.sino = plan->d_sino; a.vol = d_volume; a.tab = plan->d_tab; a.mask = plan->d_mask;
    a.nu = (int)q.width; a.nv = (int)q.slices; a.vpad = (int)plan->vpad; a.nk = (int)nk;
    a.nz4 = (int)((q.slices / 4) * 4);
    a.bstride = q.multiprojbp ? a.nz4 : (int)plan->vpad;
    a.accumulate = accumulate ? 1 : 0;
    a.stage64 = plan->stage64;
    a.tilt = plan->tilt;
    if (plan->tilt) mr::bp_tilt_terms(q.tilt_angle, q.tilt_pivot, plan->tilt_row0, a.tilt_co, a.tilt_ci);
    return mr::run_bp(a, q.interp, plan->impl, q.fused ? 0 : 1, s);
}

size_t bp_plan_slices(const mr_bp_plan *plan) { return plan->p.slices; }
void bp_set_tilt_row0(mr_bp_plan *plan, size_t row0) { plan->tilt_row0 = row0; }
size_t bp_plan_width(const mr_bp_plan *plan) { return plan->p.width; }

}  // namespace mr

extern "C" int mr_bp_parallel(mr_bp_plan *plan, const float *d_proj, size_t n_proj,
                              const float *angles, const float *weights, float *d_volume,
                              int accumulate, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(d_volume != nullptr, "volume pointer is NULL");
    MR_REQUIRE(n_proj == 0 || (d_proj && angles && weights), "NULL projection/angle/weight pointer");
    hipStream_t s = (hipStream_t)stream;
    int rc = mr::bp_load_tables(plan, n_proj, angles, weights, s);
    if (rc != MR_OK) return rc;
    rc = mr::bp_enqueue(plan, d_proj, d_volume, accumulate, s);
    if (rc != MR_OK) return rc;
    // no host wait: the next call on this plan (any stream) waits for this event
    if (!plan->done_ev) MR_HIP(hipEventCreateWithFlags(&plan->done_ev, hipEventDisableTiming));
    MR_HIP(hipEventRecord(plan->done_ev, s));
    return MR_OK;
    MR_GUARD_END
}

namespace {

// Chunks of the used projections for bp_adjoint_lds_kernel: one stepping
// class per chunk, projections in order of direction, cut where the candidate
// positions of some block of kAdjNB bins on the first or last line (the span
// of affine functions of the line is largest at an end) would spread over
// more than kAdjW - 6 voxels (floor/ceil and one voxel of margin each side).
void adj_chunks(const std::vector<float> &tab, int nu, std::vector<int4> &chunks, std::vector<int> &order)
{
    const int nk = (int)(tab.size() / 4);
    const int nbt = (nu + mr::kAdjNB - 1) / mr::kAdjNB;
    for (int cls = 0; cls < 2; ++cls) {
        std::vector<std::pair<double, int>> ks;
        for (int k = 0; k < nk; ++k) {
            const float c = tab[4 * k + 1], sn = tab[4 * k + 2];
            if ((std::fabs(sn) >= std::fabs(c) ? 0 : 1) != cls) continue;
            double dir = std::atan2((double)sn, (double)c);
            if (dir < 0) dir += 2.0 * 3.14159265358979323846;
            ks.push_back({dir, k});
        }
        std::sort(ks.begin(), ks.end());
        std::vector<double> mn(2 * nbt), mx(2 * nbt), tmn(2 * nbt), tmx(2 * nbt);
        int first = (int)order.size(), count = 0;
        auto extents = [&](int k, int bt, int end, double &lo, double &hi) {
            const double su = tab[4 * k], c = tab[4 * k + 1], sn = tab[4 * k + 2];
            const int n0 = bt * mr::kAdjNB, n1 = std::min(n0 + mr::kAdjNB - 1, nu - 1);
            const double ulo = n0 <= 1 ? -1.0 : (double)n0 - 1.0, uhi = (double)n1 + 1.0;
            const double L = end == 0 ? 1.0 : (double)nu;
            double a, b;
            if (cls == 0) {
                a = (su + c * L - uhi) / sn;
                b = (su + c * L - ulo) / sn;
            } else {
                a = (ulo + sn * L - su) / c;
                b = (uhi + sn * L - su) / c;
            }
            lo = std::min(a, b);
            hi = std::max(a, b);
        };
        for (const auto &e : ks) {
            const int k = e.second;
            double worst = 0.0;
            for (int bt = 0; bt < nbt; ++bt)
                for (int end = 0; end < 2; ++end) {
                    double lo, hi;
                    extents(k, bt, end, lo, hi);
                    const int i = 2 * bt + end;
                    tmn[i] = count ? std::min(mn[i], lo) : lo;
                    tmx[i] = count ? std::max(mx[i], hi) : hi;
                    worst = std::max(worst, tmx[i] - tmn[i]);
                }
            if (count > 0 && (count == mr::kAdjKA || worst > (double)(mr::kAdjW - 6))) {
                chunks.push_back(make_int4(first, count, cls, 0));
                first = (int)order.size();
                count = 0;
                for (int bt = 0; bt < nbt; ++bt)
                    for (int end = 0; end < 2; ++end) extents(k, bt, end, tmn[2 * bt + end], tmx[2 * bt + end]);
            }
            mn = tmn;
            mx = tmx;
            order.push_back(k);
            ++count;
        }
        if (count > 0) chunks.push_back(make_int4(first, count, cls, 0));
    }
}

}  // namespace

extern "C" int mr_bp_parallel_adjoint(mr_bp_plan *plan, const float *d_grad_volume, size_t n_proj,
                                      const float *angles, const float *weights, float *d_grad_proj, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(d_grad_volume != nullptr && d_grad_proj != nullptr, "NULL gradient pointer");
    MR_REQUIRE(n_proj == 0 || (angles && weights), "NULL angle/weight pointer");
    const mr_bp_params &q = plan->p;
    MR_REQUIRE(!(q.multiprojbp && (q.slices & 3)), "no adjoint for MultiProjBP's column offset with slices % 4 != 0");
    if (q.correct_tilt) {
        mr::set_error("no adjoint for the tilt-corrected back-projection");
        return MR_ERR_UNSUPPORTED;
    }
    // the kernel keeps every line's span in LDS (8 bytes per line)
    MR_REQUIRE(q.width <= 12288, "the back-projection adjoint supports widths up to 12288");
    hipStream_t s = (hipStream_t)stream;
    int rc = mr::fill_zero(d_grad_proj, n_proj * q.slices * q.width, s);
    if (rc != MR_OK) return rc;
    std::vector<size_t> used = mr::used_projections(n_proj, (size_t)q.buffer_size);
    const size_t nk = used.size();
    MR_REQUIRE(nk <= 65535, "at most 65535 projections per call");
    const size_t nz4 = (q.slices / 4) * 4;
    if (nk == 0 || nz4 == 0) {
        MR_HIP(hipStreamSynchronize(s));
        return MR_OK;
    }
    rc = ensure_capacity(plan, n_proj, nk);
    if (rc != MR_OK) return rc;
    if (plan->done_ev) MR_HIP(hipStreamWaitEvent(s, plan->done_ev, 0));
    plan->loaded = false;   // the tables below replace the forward ones
    std::vector<float> tab;
    mr::bp_tables(angles, used.data(), nk, q.direction, q.center, q.width / 2, tab);
    std::vector<int> idx(nk);
    for (size_t k = 0; k < nk; ++k) idx[k] = (int)used[k];
    MR_HIP(hipMemcpyAsync(plan->d_tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, s));
    MR_HIP(hipMemcpyAsync(plan->d_idx, idx.data(), sizeof(int) * nk, hipMemcpyHostToDevice, s));
    MR_HIP(hipMemcpyAsync(plan->d_w, weights, sizeof(float) * n_proj, hipMemcpyHostToDevice, s));
    if (!plan->d_cmask) {
        // per column x the rows [y0, y1) whose mask holds x: the circle mask,
        // clipped to a rectangle, is convex, so they form one run
        std::vector<int2> cm(q.width, make_int2(0x7fffffff, -1));
        std::vector<int> count(q.width, 0);
        for (size_t y = 0; y < q.width; ++y)
            for (size_t x = plan->first[y]; x < std::max(plan->first[y], plan->second[y]); ++x) {
                cm[x].x = std::min(cm[x].x, (int)y);
                cm[x].y = std::max(cm[x].y, (int)y + 1);
                ++count[x];
            }
        for (size_t x = 0; x < q.width; ++x) {
            if (count[x] == 0) cm[x] = make_int2(0, 0);
            if (count[x] != cm[x].y - cm[x].x) {
                mr::set_error("mr_bp_parallel_adjoint: a mask column is not one run");
                return MR_ERR_INTERNAL;
            }
        }
        MR_HIP(hipMalloc(&plan->d_cmask, sizeof(int2) * q.width));
        MR_HIP(hipMemcpy(plan->d_cmask, cm.data(), sizeof(int2) * q.width, hipMemcpyHostToDevice));
    }
    std::vector<int4> chunks;
    std::vector<int> order;
    adj_chunks(tab, (int)q.width, chunks, order);
    const size_t bytes = sizeof(int4) * chunks.size() + sizeof(int) * order.size();
    if (bytes > plan->adj_cap) {
        if (plan->d_adj) (void)hipFree(plan->d_adj);
        plan->d_adj = nullptr;
        plan->adj_cap = 0;
        MR_HIP(hipMalloc(&plan->d_adj, bytes));
        plan->adj_cap = bytes;
    }
    int4 *d_chunks = static_cast<int4 *>(plan->d_adj);
    int *d_order = reinterpret_cast<int *>(d_chunks + chunks.size());
    MR_HIP(hipMemcpyAsync(d_chunks, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice, s));
    MR_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, s));
    // slices per block: the largest of 32, 16, 8, 4 that divides 4*(slices/4)
    // (64 measured slower at 1024^2 x 64 x 900: 13.0 against 12.0 ms, two
    // waves per SIMD instead of three)
    const int zt = nz4 % 32 == 0 ? 32 : (nz4 % 16 == 0 ? 16 : (nz4 % 8 == 0 ? 8 : 4));
    const int nzs = (int)(nz4 / zt), nbt = (int)((q.width + mr::kAdjNB - 1) / mr::kAdjNB);
    const long total = (long)chunks.size() * nbt * nzs;
    MR_REQUIRE(total < (1L << 31) - 8, "adjoint grid too large");
    const dim3 grid((unsigned)(8 * ((total + 7) / 8)));
    const bool nn = q.interp == MR_INTERP_NEAREST;
    const size_t dyn = sizeof(int2) * q.width;
    mr::ProfScope prof("bp_adjoint", s);
    if (!nn) {
        // bins 1 and nu - 1 on the plan's side stream, beside the main kernel
        if (!plan->side) {
            MR_HIP(hipStreamCreateWithFlags(&plan->side, hipStreamNonBlocking));
            MR_HIP(hipEventCreateWithFlags(&plan->ev_fork, hipEventDisableTiming));
            MR_HIP(hipEventCreateWithFlags(&plan->ev_join, hipEventDisableTiming));
        }
        MR_HIP(hipEventRecord(plan->ev_fork, s));
        MR_HIP(hipStreamWaitEvent(plan->side, plan->ev_fork, 0));
        hipLaunchKernelGGL(mr::bp_adjoint_edge_kernel, dim3((unsigned)(2 * nk)), dim3(64), 0, plan->side,
                           d_grad_volume, d_grad_proj, plan->d_tab, plan->d_mask, plan->d_cmask, plan->d_idx,
                           plan->d_w, (int)q.width, (int)q.slices, (int)nz4);
        MR_HIP(hipGetLastError());
        MR_HIP(hipEventRecord(plan->ev_join, plan->side));
    }
#define MR_ADJ_LAUNCH(I, Z)                                                                                        \
    hipLaunchKernelGGL((mr::bp_adjoint_kernel<I, Z>), grid, dim3(256), dyn, s, d_grad_volume, d_grad_proj,        \
                       plan->d_tab, plan->d_mask, plan->d_cmask, plan->d_idx, plan->d_w, d_chunks, d_order,      \
                       (int)chunks.size(), nbt, nzs, (int)q.width, (int)q.slices)
    switch (zt) {
    case 32: if (nn) MR_ADJ_LAUNCH(1, 32); else MR_ADJ_LAUNCH(0, 32); break;
    case 16: if (nn) MR_ADJ_LAUNCH(1, 16); else MR_ADJ_LAUNCH(0, 16); break;
    case 8: if (nn) MR_ADJ_LAUNCH(1, 8); else MR_ADJ_LAUNCH(0, 8); break;
    default: if (nn) MR_ADJ_LAUNCH(1, 4); else MR_ADJ_LAUNCH(0, 4); break;
    }
#undef MR_ADJ_LAUNCH
    MR_HIP(hipGetLastError());
    if (!nn) MR_HIP(hipStreamWaitEvent(s, plan->ev_join, 0));
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fbp_parallel(mr_filter_plan *fplan, mr_bp_plan *bplan, float *d_proj,
                               size_t n_proj, const float *angles, const float *weights,
                               float *d_volume, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(fplan != nullptr && bplan != nullptr, "plan is NULL");
    int rc = mr_filter_rows(fplan, d_proj, n_proj * bplan->p.slices, stream);
    if (rc != MR_OK) return rc;
    return mr_bp_parallel(bplan, d_proj, n_proj, angles, weights, d_volume, 0, stream);
    MR_GUARD_END
}

// --------------------------------------------------------------------------
// Batched back-projection of many uneven stacks (C ABI mr_batch_*).
// Per stack: ProjectionFilterSingle (out of place, the raw stack is kept),
// weighting/transposition into the stack's sinogram, then ONE persistent
// launch over the work items of all stacks.  mr_batch_run only enqueues
// kernels (no host synchronisation, no host copies), so it can be captured
// in a hipGraph; mr_batch_run_graph does that on its first call.
// --------------------------------------------------------------------------
#include "mr_filter.h"

namespace {

struct BatchStack {
    mr_batch_stack desc{};
    mr_filter_plan *fplan = nullptr;
    float *d_filt = nullptr;   // filtered copy of the raw stack
    float *d_sino = nullptr;   // [n_used][width][vpad]
    float4 *d_tab = nullptr;
    int2 *d_mask = nullptr;
    int *d_idx = nullptr;
    float *d_w = nullptr;
    size_t vpad = 0, n_used = 0;
    mr::BpArgs args{};
};

struct Variant {
    int zt, kc;
};
const Variant kVariants[] = {{32, 4}, {32, 8}, {16, 8}, {16, 4}, {16, 16}};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
// item list index of a slice tile (lists for ZT = 32 and 16)
inline int zt_list(int zt) { return zt == 32 ? 0 : 1; }

}  // namespace

struct mr_batch_plan {
    std::vector<BatchStack> stacks;
    mr::BpArgs *d_args = nullptr;
    int4 *d_items[2] = {nullptr, nullptr};   // item lists for ZT = 32 and ZT = 16
    int *d_parts[2] = {nullptr, nullptr};    // 9 partition boundaries per list
    int n_items[2] = {0, 0};
    int *d_counter = nullptr;                // 8 partition counters
    int variant = 0;
    int exact = 1;                           // 0: interpolation folded into FMAs (MR_BATCH_FUSED)
    int grid[kNumVariants] = {0};
    double updates = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

int batch_launch_bp(mr_batch_plan *p, int variant, hipStream_t s)
{
    const Variant v = kVariants[variant];
    const int li = zt_list(v.zt);
    hipLaunchKernelGGL(mr::reset_counter_kernel, dim3(1), dim3(64), 0, s, p->d_counter);
    MR_HIP(hipGetLastError());
    mr::ProfScope prof("bp_batch", s);
    dim3 grid((unsigned)p->grid[variant]), block(256);
#define MR_BATCH_CASE(ZT, KC)                                                                         \
    if (v.zt == ZT && v.kc == KC) {                                                                   \
        if (p->exact)                                                                                 \
            hipLaunchKernelGGL((mr::bp_batch_kernel<ZT, 0, true, KC>), grid, block, 0, s, p->d_args,   \
                               p->d_items[li], p->d_parts[li], p->d_counter);                         \
        else                                                                                          \
            hipLaunchKernelGGL((mr::bp_batch_kernel<ZT, 0, false, KC>), grid, block, 0, s, p->d_args,  \
                               p->d_items[li], p->d_parts[li], p->d_counter);                         \
    }
    MR_BATCH_CASE(32, 4)
    MR_BATCH_CASE(32, 8)
    MR_BATCH_CASE(16, 8)
    MR_BATCH_CASE(16, 4)
    MR_BATCH_CASE(16, 16)
#undef MR_BATCH_CASE
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int batch_prepare(mr_batch_plan *p, hipStream_t s)
{
    for (auto &b : p->stacks) {
        const mr_batch_stack &d = b.desc;
        int rc = mr::filter_rows_to(b.fplan, d.d_proj, b.d_filt, d.n_proj * d.slices, s);
        if (rc != MR_OK) return rc;
        if (b.n_used == 0) continue;
        rc = mr::launch_bp_prep(b.d_filt, b.d_sino, b.d_idx, b.d_w, (int)d.width, (int)d.slices, (int)b.vpad,
                                (int)b.n_used, s);
        if (rc != MR_OK) return rc;
    }
    return MR_OK;
}

void batch_free(mr_batch_plan *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (auto &b : p->stacks) {
        if (b.fplan) mr_filter_plan_destroy(b.fplan);
        for (void *ptr : {(void *)b.d_filt, (void *)b.d_sino, (void *)b.d_tab, (void *)b.d_mask,
                          (void *)b.d_idx, (void *)b.d_w})
            if (ptr) (void)hipFree(ptr);
    }
    for (void *ptr : {(void *)p->d_args, (void *)p->d_items[0], (void *)p->d_items[1], (void *)p->d_parts[0],
                      (void *)p->d_parts[1], (void *)p->d_counter})
        if (ptr) (void)hipFree(ptr);
    delete p;
}

template <int ZT, int KC>
int occupancy_grid(int &grid, int exact)
{
    int dev = 0, ncu = 0, nb = 0;
    MR_HIP(hipGetDevice(&dev));
    MR_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (exact)
        MR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mr::bp_batch_kernel<ZT, 0, true, KC>, 256, 0));
    else
        MR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mr::bp_batch_kernel<ZT, 0, false, KC>, 256, 0));
    grid = std::max(1, nb) * std::max(1, ncu);
    return MR_OK;
}

}  // namespace

extern "C" int mr_batch_plan_create(mr_batch_plan **plan, const mr_batch_stack *stacks, size_t n_stacks,
                                    int flags)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && stacks != nullptr && n_stacks > 0, "NULL argument or empty batch");
    MR_REQUIRE((flags & ~(MR_BATCH_TUNE | MR_BATCH_FUSED)) == 0, "unknown batch flags");
    *plan = nullptr;
    const bool tune = (flags & MR_BATCH_TUNE) != 0;
    auto *p = new mr_batch_plan();
    p->exact = (flags & MR_BATCH_FUSED) ? 0 : 1;
    p->stacks.resize(n_stacks);
    std::vector<std::pair<double, int4>> items[2];   // ZT = 32, ZT = 16
    for (size_t i = 0; i < n_stacks; ++i) {
        BatchStack &b = p->stacks[i];
        b.desc = stacks[i];
        const mr_batch_stack &d = b.desc;
        if (!(d.width >= 2 && d.slices >= 1 && d.n_proj >= 1 && d.d_proj && d.d_volume && d.angles &&
              d.weights) || d.direction < 0 || d.direction > 1) {
            batch_free(p);
            mr::set_error("invalid stack " + std::to_string(i));
            return MR_ERR_INVALID;
        }
        int rc = mr_filter_plan_create(&b.fplan, d.width, d.filter_type, d.cutoff, d.order, d.use_bias,
                                       d.bias_weight, d.padding_doubler);
        if (rc != MR_OK) {
            batch_free(p);
            return rc;
        }
        b.vpad = (d.slices + 3) & ~(size_t)3;
        std::vector<size_t> used = mr::used_projections(d.n_proj, (size_t)std::max(0, d.buffer_size));
        b.n_used = used.size();
        std::vector<float> tab;
        mr::bp_tables(d.angles, used.data(), b.n_used, d.direction, d.center, d.width / 2, tab);
        std::vector<int> idx(b.n_used);
        for (size_t k = 0; k < b.n_used; ++k) idx[k] = (int)used[k];
        std::vector<size_t> first, second;
        mr::circle_mask(d.width, d.width, false, nullptr, false, first, second);
        std::vector<int2> m(d.width);
        for (size_t y = 0; y < d.width; ++y) {
            m[y].x = (int)first[y];
            m[y].y = (int)std::max(first[y], second[y]);
        }
        hipError_t e = hipMalloc(&b.d_filt, sizeof(float) * d.n_proj * d.slices * d.width);
        if (e == hipSuccess) e = hipMalloc(&b.d_sino, sizeof(float) * std::max<size_t>(b.n_used, 1) * d.width * b.vpad);
        if (e == hipSuccess) e = hipMalloc(&b.d_tab, sizeof(float4) * std::max<size_t>(b.n_used, 1));
        if (e == hipSuccess) e = hipMalloc(&b.d_mask, sizeof(int2) * d.width);
        if (e == hipSuccess) e = hipMalloc(&b.d_idx, sizeof(int) * std::max<size_t>(b.n_used, 1));
        if (e == hipSuccess) e = hipMalloc(&b.d_w, sizeof(float) * d.n_proj);
        if (e == hipSuccess && b.n_used)
            e = hipMemcpy(b.d_tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b.d_mask, m.data(), sizeof(int2) * d.width, hipMemcpyHostToDevice);
        if (e == hipSuccess && b.n_used)
            e = hipMemcpy(b.d_idx, idx.data(), sizeof(int) * b.n_used, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b.d_w, d.weights, sizeof(float) * d.n_proj, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            batch_free(p);
            return mr::hip_fail(e, "mr_batch_plan_create", __FILE__, __LINE__);
        }
        mr::BpArgs &a = b.args;
        a.sino = b.d_sino; a.vol = d.d_volume; a.tab = b.d_tab; a.mask = b.d_mask;
        a.nu = (int)d.width; a.nv = (int)d.slices; a.vpad = (int)b.vpad; a.nk = (int)b.n_used;
        a.nz4 = (int)((d.slices / 4) * 4); a.bstride = (int)b.vpad; a.accumulate = 0;
        p->updates += (double)d.width * d.width * d.slices * d.n_proj;
        // work items in spatial order (slice chunk, stripes of 4 tile
        // columns, rows, columns), cost ~ projections x voxels inside the circle
        const int tiles = (int)((d.width + 15) / 16);
        for (int li = 0; li < 2; ++li) {
            const int ZT = li == 0 ? 32 : 16;
            const int zc = (int)((d.slices + ZT - 1) / ZT);
            for (int bz = 0; bz < zc; ++bz)
                for (int sx = 0; sx < tiles; sx += 4)
                    for (int by = 0; by < tiles; ++by)
                        for (int bx = sx; bx < std::min(tiles, sx + 4); ++bx) {
                            double inside = 0;
                            for (int yy = by * 16; yy < std::min<int>((int)d.width, by * 16 + 16); ++yy) {
                                const int lo = std::max(m[yy].x, bx * 16), hi = std::min(m[yy].y, bx * 16 + 16);
                                inside += std::max(0, hi - lo);
                            }
                            const double cost = 1.0 + inside * (double)b.n_used * std::min<int>(ZT, (int)d.slices - bz * ZT);
                            items[li].push_back({cost, make_int4((int)i, bx, by, bz)});
                        }
        }
    }
    for (int li = 0; li < 2; ++li) {
        // eight contiguous partitions of equal cost, one per XCD
        const auto &v = items[li];
        double total = 0;
        for (const auto &it : v) total += it.first;
        std::vector<int> parts(9, (int)v.size());
        parts[0] = 0;
        double acc = 0;
        int next = 1;
        for (size_t k = 0; k < v.size() && next < 8; ++k) {
            acc += v[k].first;
            while (next < 8 && acc >= total * next / 8.0) parts[next++] = (int)k + 1;
        }
        std::vector<int4> flat(v.size());
        for (size_t k = 0; k < v.size(); ++k) flat[k] = v[k].second;
        p->n_items[li] = (int)flat.size();
        hipError_t e = hipMalloc(&p->d_items[li], sizeof(int4) * std::max<size_t>(flat.size(), 1));
        if (e == hipSuccess && !flat.empty())
            e = hipMemcpy(p->d_items[li], flat.data(), sizeof(int4) * flat.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&p->d_parts[li], sizeof(int) * 9);
        if (e == hipSuccess) e = hipMemcpy(p->d_parts[li], parts.data(), sizeof(int) * 9, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            batch_free(p);
            return mr::hip_fail(e, "mr_batch_plan_create items", __FILE__, __LINE__);
        }
    }
    std::vector<mr::BpArgs> args(n_stacks);
    for (size_t i = 0; i < n_stacks; ++i) args[i] = p->stacks[i].args;
    hipError_t e = hipMalloc(&p->d_args, sizeof(mr::BpArgs) * n_stacks);
    if (e == hipSuccess) e = hipMemcpy(p->d_args, args.data(), sizeof(mr::BpArgs) * n_stacks, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_counter, sizeof(int) * 8);
    if (e != hipSuccess) {
        batch_free(p);
        return mr::hip_fail(e, "mr_batch_plan_create args", __FILE__, __LINE__);
    }
    int rc = MR_OK;
    if (rc == MR_OK) rc = occupancy_grid<32, 4>(p->grid[0], p->exact);
    if (rc == MR_OK) rc = occupancy_grid<32, 8>(p->grid[1], p->exact);
    if (rc == MR_OK) rc = occupancy_grid<16, 8>(p->grid[2], p->exact);
    if (rc == MR_OK) rc = occupancy_grid<16, 4>(p->grid[3], p->exact);
    if (rc == MR_OK) rc = occupancy_grid<16, 16>(p->grid[4], p->exact);
    if (rc != MR_OK) {
        batch_free(p);
        return rc;
    }
    if (tune) {
        // Autotuner: time each (slice tile, projection chunk) variant on the
        // real batch and keep the fastest.
        hipStream_t s = nullptr;
        hipEvent_t a0, a1;
        MR_HIP(hipEventCreate(&a0));
        MR_HIP(hipEventCreate(&a1));
        rc = batch_prepare(p, s);
        float best = 1e30f;
        for (int v = 0; v < kNumVariants && rc == MR_OK; ++v) {
            rc = batch_launch_bp(p, v, s);  // warm-up
            if (rc != MR_OK) break;
            MR_HIP(hipEventRecord(a0, s));
            rc = batch_launch_bp(p, v, s);
            MR_HIP(hipEventRecord(a1, s));
            MR_HIP(hipEventSynchronize(a1));
            float ms = 0.0f;
            MR_HIP(hipEventElapsedTime(&ms, a0, a1));
            if (ms < best) {
                best = ms;
                p->variant = v;
            }
        }
        (void)hipEventDestroy(a0);
        (void)hipEventDestroy(a1);
        if (rc != MR_OK) {
            batch_free(p);
            return rc;
        }
    }
    *plan = p;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_batch_plan_info(const mr_batch_plan *plan, int *slice_tile, int *proj_chunk,
                                  size_t *n_items, double *updates)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    const Variant v = kVariants[plan->variant];
    if (slice_tile) *slice_tile = v.zt;
    if (proj_chunk) *proj_chunk = v.kc;
    if (n_items) *n_items = (size_t)plan->n_items[zt_list(v.zt)];
    if (updates) *updates = plan->updates;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_batch_run(mr_batch_plan *plan, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    int rc = batch_prepare(plan, s);
    if (rc != MR_OK) return rc;
    return batch_launch_bp(plan, plan->variant, s);
    MR_GUARD_END
}

extern "C" int mr_batch_run_graph(mr_batch_plan *plan, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    MR_REQUIRE(s != nullptr, "graph launches need a non-default stream");
    if (!plan->exec) {
        MR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = batch_prepare(plan, s);
        if (rc == MR_OK) rc = batch_launch_bp(plan, plan->variant, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != MR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return mr::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        plan->graph = g;
        MR_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
    }
    mr::ProfScope prof("bp_batch_graph", s);
    MR_HIP(hipGraphLaunch(plan->exec, s));
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_batch_run_sequential(mr_batch_plan *plan, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    int rc = batch_prepare(plan, s);
    if (rc != MR_OK) return rc;
    for (auto &b : plan->stacks) {
        rc = mr::run_bp(b.args, 0, 1, plan->exact, s);
        if (rc != MR_OK) return rc;
    }
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_batch_plan_destroy(mr_batch_plan *plan)
{
    MR_GUARD_BEGIN
    if (plan) batch_free(plan);
    return MR_OK;
    MR_GUARD_END
}
