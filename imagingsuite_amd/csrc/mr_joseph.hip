// Matched Joseph projector pair for parallel beam on gfx950: forward A and
// its exact adjoint A^T.  No reference counterpart exists (the reference's
// LinearForwardProjector, frameworks/tomography/Framework/ReconAlgorithms/
// ReconAlgorithms/linearforwardprojector.cpp:15-143, is 2-D only and is not
// the adjoint of any back-projector); the geometry is the one of the
// reference's parallel-beam back-projector (StdBackProjectorBase.cpp:73-80):
//   u(x, y) = center + cos(phi) (y + 1 - c) - sin(phi) (x + 1 - c),  c = nx / 2.
//
// With m = max(|sin phi|, |cos phi|) both operators use the same kernel
//   K(t) = max(0, 1 - |t| / m) / m,
// A[phi][n][z]  = sum_{x,y} K(u(x,y) - n) vol[y][x][z]   (ray-driven Joseph
//                 stepping along the axis the ray is closest to: linear
//                 interpolation between the two nearest voxels, step 1/m)
// AT[y][x][z]   = sum_{phi,n} K(u(x,y) - n) sino[phi][n][z]   (voxel-driven:
//                 the two detector bins around u with weights K)
// so A^T is the transpose of A in exact arithmetic.
//
// Layout: volume [nx(y)][nx(x)][nz] (MatrixZXY), sinogram [n_ang][n_det][nz]
// (slice fastest).  Each thread owns ZT slices held as float4 vectors.
#include "mr_common.h"
#include "mr_filter.h"
#include "mr_joseph.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace mr {

struct JosephArgs {
    const float *src;
    float *dst;
    const float4 *tab;  // per angle: (su, cos, sin, 1/m)   su = c (sin - cos) + center
    int nx, nz, nd, na;
};

// Loads below are never predicated: out-of-volume neighbours read a clamped,
// valid address with weight zero, so the compiler can keep the loads of
// several unrolled steps in flight together.
template <int ZT>
__device__ __forceinline__ void lerp_accumulate(float (&acc)[ZT], const float *p0, const float *p1, float w0,
                                                float w1)
{
#pragma unroll
    for (int g = 0; g < ZT / 4; ++g) {
        const float4 v0 = *reinterpret_cast<const float4 *>(p0 + 4 * g);
        const float4 v1 = *reinterpret_cast<const float4 *>(p1 + 4 * g);
        acc[4 * g + 0] = __builtin_fmaf(w1, v1.x, __builtin_fmaf(w0, v0.x, acc[4 * g + 0]));
        acc[4 * g + 1] = __builtin_fmaf(w1, v1.y, __builtin_fmaf(w0, v0.y, acc[4 * g + 1]));
        acc[4 * g + 2] = __builtin_fmaf(w1, v1.z, __builtin_fmaf(w0, v0.z, acc[4 * g + 2]));
        acc[4 * g + 3] = __builtin_fmaf(w1, v1.w, __builtin_fmaf(w0, v0.w, acc[4 * g + 3]));
    }
}

// Range [lo, hi) of the step index t = 0..N-1 for which the crossing
// position q(t) = A + B t can fall in [-1, N) (where a voxel pair gets a
// nonzero weight), widened by two steps against rounding; the loop body
// still tests every step exactly.
__device__ __forceinline__ void step_range(float A, float B, int N, int &lo, int &hi)
{
    lo = 0;
    hi = N;
    if (fabsf(B) > 1e-20f) {
        float t1 = (-1.0f - A) / B, t2 = ((float)N - A) / B;
        if (t1 > t2) {
            const float tmp = t1;
            t1 = t2;
            t2 = tmp;
        }
        lo = max(0, (int)fmaxf(floorf(t1) - 2.0f, -1.0f));
        hi = min(N, (int)fminf(ceilf(t2) + 2.0f, (float)N));
    } else if (A < -1.0f || A >= (float)N) {
        hi = 0;
    }
}

// Forward: one thread per (angle, bin) and ZT slices.
template <int ZT>
__global__ __launch_bounds__(256) void joseph_forward_kernel(JosephArgs a)
{
    const int n = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ang = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * ZT;
    if (n >= a.nd || ang >= a.na) return;
    const float4 t = a.tab[ang];
    const float su = t.x, c = t.y, s = t.z, inv_m = t.w;
    const int N = a.nx;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    const float fn = (float)n;
    if (fabsf(s) >= fabsf(c)) {
        // step over rows y: x(y) solves u(x, y) = n
        const float inv_s = 1.0f / s;
        int lo, hi;
        step_range((su + c - fn) * inv_s - 1.0f, c * inv_s, N, lo, hi);
#pragma unroll 2
        for (int y = lo; y < hi; ++y) {
            const float xf = (su + c * (float)(y + 1) - fn) * inv_s - 1.0f;
            const float fl = floorf(xf);
            const int x0 = (int)fl;
            const float f = xf - fl;
            const float w0 = (x0 >= 0 && x0 < N) ? 1.0f - f : 0.0f;
            const float w1 = (x0 + 1 >= 0 && x0 + 1 < N) ? f : 0.0f;
            const int xa = min(max(x0, 0), N - 1), xb = min(max(x0 + 1, 0), N - 1);
            const float *row = a.src + ((size_t)y * N) * a.nz + z0;
            lerp_accumulate<ZT>(acc, row + (size_t)xa * a.nz, row + (size_t)xb * a.nz, w0, w1);
        }
    } else {
        // step over columns x: y(x) solves u(x, y) = n
        const float inv_c = 1.0f / c;
        int lo, hi;
        step_range((fn - su + s) * inv_c - 1.0f, s * inv_c, N, lo, hi);
#pragma unroll 2
        for (int x = lo; x < hi; ++x) {
            const float yf = (fn - su + s * (float)(x + 1)) * inv_c - 1.0f;
            const float fl = floorf(yf);
            const int y0 = (int)fl;
            const float f = yf - fl;
            const float w0 = (y0 >= 0 && y0 < N) ? 1.0f - f : 0.0f;
            const float w1 = (y0 + 1 >= 0 && y0 + 1 < N) ? f : 0.0f;
            const int ya = min(max(y0, 0), N - 1), yb = min(max(y0 + 1, 0), N - 1);
            const float *col = a.src + (size_t)x * a.nz + z0;
            lerp_accumulate<ZT>(acc, col + (size_t)ya * N * a.nz, col + (size_t)yb * N * a.nz, w0, w1);
        }
    }
    float *out = a.dst + ((size_t)ang * a.nd + n) * a.nz + z0;
#pragma unroll
    for (int g = 0; g < ZT / 4; ++g)
        *reinterpret_cast<float4 *>(out + 4 * g) =
            make_float4(acc[4 * g] * inv_m, acc[4 * g + 1] * inv_m, acc[4 * g + 2] * inv_m, acc[4 * g + 3] * inv_m);
}

// LDS-staged forward over a chunk of angles that all step along the same
// axis (class 0: |sin| >= |cos|, rows y; class 1: columns x).  A block owns
// kFwdNB detector bins x (up to) kFwdKA angles x ZT slices, one ray per
// thread.  It walks the lines L = 0..N-1 of its class R at a time: for
// each line it stages the kFwdW voxels (x ZT slices) from the lowest crossing
// of any of its rays, then every ray interpolates its two voxels from LDS.
// Staged positions outside the volume are zeros and carry weight zero, so the
// sums are those of the direct kernel, accumulated over the same lines in
// the same order.  Neighbouring angles cross a line close together, so one
// staged segment serves all the chunk's rays: the volume is read once per
// chunk instead of once per angle.  The host builds the chunks so that every
// line's crossings of a block fit in kFwdW - 2 voxels.
constexpr int kFwdNB = 16, kFwdKA = 16, kFwdR = 1, kFwdW = 80;

template <int ZT, int R>
__global__ __launch_bounds__(256) void joseph_forward_lds_kernel(JosephArgs a, const int4 *__restrict__ chunks,
                                                                 const int *__restrict__ order, int n_chunks,
                                                                 int nbt, int nzs)
{
    constexpr int NB = kFwdNB, W = kFwdW, RS = ZT + 4, G = ZT / 4;
    __shared__ float seg[R][W][RS];
    __shared__ int lo_s[R], hi_s[R];
    const int total = n_chunks * nbt * nzs;
    const int per = (total + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int wi = (b & 7) * per + (b >> 3);   // XCD b % 8 takes the (b % 8)-th eighth
    if (wi >= total) return;
    const int bt = wi % nbt;
    const int rest = wi / nbt;
    const int ci = rest % n_chunks;
    const int zs = rest / n_chunks;
    const int4 ch = chunks[ci];                // first, count, class
    const int ai = threadIdx.x / NB, bi = threadIdx.x % NB;
    const int n = bt * NB + bi;
    const bool active = ai < ch.y && n < a.nd;
    const bool edge = active && (bi == 0 || bi == NB - 1 || n == a.nd - 1);
    const int ang = order[ch.x + min(ai, ch.y - 1)];
    const float4 t = a.tab[ang];
    const float su = t.x, c = t.y, s = t.z, inv_m = t.w;
    const int N = a.nx;
    const int z0 = zs * ZT;
    const bool rows = ch.z == 0;
    // crossing of line L: ((a0 + b1 (L+1)) + cn) * inv - 1
    const float a0 = rows ? su : -su, b1 = rows ? c : s;
    const float fn = (float)n;
    const float cn = rows ? -fn : fn;
    const float inv = rows ? 1.0f / s : 1.0f / c;
    const size_t strideL = rows ? (size_t)N * a.nz : (size_t)a.nz;
    const size_t strideP = rows ? (size_t)a.nz : (size_t)N * a.nz;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    for (int L0 = 0; L0 < N; L0 += R) {
        if (threadIdx.x < R) {
            lo_s[threadIdx.x] = 0x7fffffff;
            hi_s[threadIdx.x] = -0x7fffffff;
        }
        __syncthreads();
        if (edge) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float pf = ((a0 + b1 * (float)(L0 + r + 1)) + cn) * inv - 1.0f;
                const int p0 = (int)floorf(pf);
                atomicMin(&lo_s[r], p0);
                atomicMax(&hi_s[r], p0);
            }
        }
        __syncthreads();
        // the group is skipped when no ray crosses the volume on its lines
        bool any = false;
#pragma unroll
        for (int r = 0; r < R; ++r)
            any = any || (L0 + r < N && hi_s[r] >= -1 && lo_s[r] < N);
        if (any) {
            // only the voxels a line's rays read: lo .. hi + 1
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int L = L0 + r, lo = lo_s[r], nj = min(hi_s[r] - lo + 2, W);
                const float *line = a.src + (size_t)min(L, N - 1) * strideL + z0;
                for (int e = threadIdx.x; e < nj * G; e += 256) {
                    const int j = e / G, g = e - j * G;
                    const int p = lo + j;
                    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (L < N && p >= 0 && p < N)
                        v = *reinterpret_cast<const float4 *>(line + (size_t)p * strideP + 4 * g);
                    *reinterpret_cast<float4 *>(&seg[r][j][4 * g]) = v;
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (L0 + r < N) {
                    const float pf = ((a0 + b1 * (float)(L0 + r + 1)) + cn) * inv - 1.0f;
                    const float fl = floorf(pf);
                    const int p0 = (int)fl;
                    const float f = pf - fl;
                    const float w0 = (p0 >= 0 && p0 < N) ? 1.0f - f : 0.0f;
                    const float w1 = (p0 + 1 >= 0 && p0 + 1 < N) ? f : 0.0f;
                    const int j = min(max(p0 - lo_s[r], 0), W - 2);
                    const float *v0 = &seg[r][j][0];
                    const float *v1 = v0 + RS;
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const float4 x0 = *reinterpret_cast<const float4 *>(v0 + 4 * g);
                        const float4 x1 = *reinterpret_cast<const float4 *>(v1 + 4 * g);
                        acc[4 * g + 0] = __builtin_fmaf(w1, x1.x, __builtin_fmaf(w0, x0.x, acc[4 * g + 0]));
                        acc[4 * g + 1] = __builtin_fmaf(w1, x1.y, __builtin_fmaf(w0, x0.y, acc[4 * g + 1]));
                        acc[4 * g + 2] = __builtin_fmaf(w1, x1.z, __builtin_fmaf(w0, x0.z, acc[4 * g + 2]));
                        acc[4 * g + 3] = __builtin_fmaf(w1, x1.w, __builtin_fmaf(w0, x0.w, acc[4 * g + 3]));
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    float *out = a.dst + ((size_t)ang * a.nd + n) * a.nz + z0;
#pragma unroll
    for (int g = 0; g < G; ++g)
        *reinterpret_cast<float4 *>(out + 4 * g) =
            make_float4(acc[4 * g] * inv_m, acc[4 * g + 1] * inv_m, acc[4 * g + 2] * inv_m, acc[4 * g + 3] * inv_m);
}

// The same forward with one barrier per line.  The block first computes,
// for every line, the span of its rays' crossings (the minimum and maximum
// floor(crossing) over the edge rays, the first and last valid bin of each
// of its angles, with the same float expression) into LDS; the per-line
// loop then stages line L + 1 into one buffer while its rays interpolate
// line L from the other, and a single barrier separates the lines.  Staged
// values, offsets and the order of the sums are those of
// joseph_forward_lds_kernel, so the projections are bit-identical to it.
// Dynamic LDS: 2 N ints (the spans).  Bounded to five waves per SIMD, which
// the LDS allows (1024^2 x 64 x 900: 6.8 ms, 7.7 ms at the four the
// registers would otherwise give).
template <int ZT>
__global__ __launch_bounds__(256, 5) void joseph_forward_lds2_kernel(JosephArgs a, const int4 *__restrict__ chunks,
                                                                  const int *__restrict__ order, int n_chunks,
                                                                  int nbt, int nzs)
{
    constexpr int NB = kFwdNB, KA = kFwdKA, W = kFwdW, RS = ZT + 4, G = ZT / 4;
    __shared__ float seg[2][W][RS];
    __shared__ float4 ep[2 * KA];   // edge rays: (a0, b1, cn, inv)
    extern __shared__ int spans[];  // lo of line L at [L], hi at [N + L]
    const int total = n_chunks * nbt * nzs;
    const int per = (total + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int wi = (b & 7) * per + (b >> 3);   // XCD b % 8 takes the (b % 8)-th eighth
    if (wi >= total) return;
    const int bt = wi % nbt;
    const int rest = wi / nbt;
    const int ci = rest % n_chunks;
    const int zs = rest / n_chunks;
    const int4 ch = chunks[ci];                // first, count, class
    const int tid = (int)threadIdx.x;
    // lanes 16 q .. 16 q + 15 of a wave take 4 angles x 4 neighbouring bins,
    // whose crossings of a line lie within a few staged rows: one angle's 16
    // bins spread over about 20 rows, and rows 16 apart share LDS banks
    const int ai = ((tid >> 2) & 3) | ((tid >> 4) & 12), bi = (tid & 3) | ((tid >> 2) & 12);
    const int n = bt * NB + bi;
    const bool active = ai < ch.y && n < a.nd;
    const int ang = order[ch.x + min(ai, ch.y - 1)];
    const float4 t = a.tab[ang];
    const float su = t.x, c = t.y, s = t.z, inv_m = t.w;
    const int N = a.nx;
    const int z0 = zs * ZT;
    const bool rows = ch.z == 0;
    // crossing of line L: ((a0 + b1 (L+1)) + cn) * inv - 1
    const float a0 = rows ? su : -su, b1 = rows ? c : s;
    const float fn = (float)n;
    const float cn = rows ? -fn : fn;
    const float inv = rows ? 1.0f / s : 1.0f / c;
    const size_t strideL = rows ? (size_t)N * a.nz : (size_t)a.nz;
    const size_t strideP = rows ? (size_t)a.nz : (size_t)N * a.nz;
    if (tid < 2 * KA) {
        // edge ray tid: angle tid / 2, the block's first (even) or last valid (odd) bin
        const int ea = tid >> 1;
        const int en = min(bt * NB + ((tid & 1) ? NB - 1 : 0), a.nd - 1);
        const float4 et = a.tab[order[ch.x + min(ea, ch.y - 1)]];
        const float efn = (float)en;
        ep[tid] = make_float4(rows ? et.x : -et.x, rows ? et.y : et.z, rows ? -efn : efn,
                              rows ? 1.0f / et.z : 1.0f / et.y);
    }
    __syncthreads();
    const int ne = 2 * ch.y;
    for (int L = tid; L < N; L += 256) {
        int lo = 0x7fffffff, hi = -0x7fffffff;
        for (int e = 0; e < ne; ++e) {
            const float4 q = ep[e];
            const int p0 = (int)floorf(((q.x + q.y * (float)(L + 1)) + q.z) * q.w - 1.0f);
            lo = min(lo, p0);
            hi = max(hi, p0);
        }
        spans[L] = lo;
        spans[N + L] = hi;
    }
    __syncthreads();
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    // a line is skipped when no ray of the block crosses the volume on it
    auto live = [&](int L) { return spans[N + L] >= -1 && spans[L] < N; };
    // only the voxels a line's rays read: lo .. hi + 1.  fetch() issues the
    // global loads of line L into registers, put() writes them to seg[buf]
    // one line later, so the loads of line L + 2 are in flight while line L
    // is consumed.  Element e = tid + 256 i is (position lo + e / G, slice
    // group e % G).
    constexpr int NE = (W * G + 255) / 256;
    float4 pre[NE];
    auto fetch = [&](int L) {
        const int lo = spans[L], nj = min(spans[N + L] - lo + 2, W);
        const float *line = a.src + (size_t)L * strideL + z0;
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int j = e / G, g = e - j * G;
            const int p = lo + j;
            pre[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (e < nj * G && p >= 0 && p < N)
                pre[i] = *reinterpret_cast<const float4 *>(line + (size_t)p * strideP + 4 * g);
        }
    };
    auto put = [&](int L, int buf) {
        const int nj = min(spans[N + L] - spans[L] + 2, W);
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int j = e / G, g = e - j * G;
            if (e < nj * G) *reinterpret_cast<float4 *>(&seg[buf][j][4 * g]) = pre[i];
        }
    };
    if (live(0)) {
        fetch(0);
        put(0, 0);
    }
    if (N > 1 && live(1)) fetch(1);
    __syncthreads();
    for (int L = 0; L < N; ++L) {
        const int buf = L & 1;
        // line L + 1: loaded during the previous line, into the other buffer
        if (L + 1 < N && live(L + 1)) put(L + 1, buf ^ 1);
        if (L + 2 < N && live(L + 2)) fetch(L + 2);
        if (live(L)) {
            const float pf = ((a0 + b1 * (float)(L + 1)) + cn) * inv - 1.0f;
            const float fl = floorf(pf);
            const int p0 = (int)fl;
            const float f = pf - fl;
            const float w0 = (p0 >= 0 && p0 < N) ? 1.0f - f : 0.0f;
            const float w1 = (p0 + 1 >= 0 && p0 + 1 < N) ? f : 0.0f;
            const int j = min(max(p0 - spans[L], 0), W - 2);
            const float *v0 = &seg[buf][j][0];
            const float *v1 = v0 + RS;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float4 x0 = *reinterpret_cast<const float4 *>(v0 + 4 * g);
                const float4 x1 = *reinterpret_cast<const float4 *>(v1 + 4 * g);
                acc[4 * g + 0] = __builtin_fmaf(w1, x1.x, __builtin_fmaf(w0, x0.x, acc[4 * g + 0]));
                acc[4 * g + 1] = __builtin_fmaf(w1, x1.y, __builtin_fmaf(w0, x0.y, acc[4 * g + 1]));
                acc[4 * g + 2] = __builtin_fmaf(w1, x1.z, __builtin_fmaf(w0, x0.z, acc[4 * g + 2]));
                acc[4 * g + 3] = __builtin_fmaf(w1, x1.w, __builtin_fmaf(w0, x0.w, acc[4 * g + 3]));
            }
        }
        __syncthreads();
    }
    if (!active) return;
    float *out = a.dst + ((size_t)ang * a.nd + n) * a.nz + z0;
#pragma unroll
    for (int g = 0; g < G; ++g)
        *reinterpret_cast<float4 *>(out + 4 * g) =
            make_float4(acc[4 * g] * inv_m, acc[4 * g + 1] * inv_m, acc[4 * g + 2] * inv_m, acc[4 * g + 3] * inv_m);
}

// Adjoint: one thread per (x, y) voxel column and ZT slices.
template <int ZT>
__global__ __launch_bounds__(256) void joseph_adjoint_kernel(JosephArgs a)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int z0 = blockIdx.z * ZT;
    const int N = a.nx;
    if (x >= N || y >= N) return;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    const float xf = (float)(x + 1), yf = (float)(y + 1);
#pragma unroll 2
    for (int ang = 0; ang < a.na; ++ang) {
        const float4 t = a.tab[ang];
        const float u = (t.x + t.y * yf) - t.z * xf;
        const float fl = floorf(u);
        const int n1 = (int)fl;
        const float f = u - fl;
        const float inv_m = t.w;
        const bool ok1 = n1 >= 0 && n1 < a.nd, ok2 = n1 + 1 >= 0 && n1 + 1 < a.nd;
        const float w1 = ok1 ? fmaxf(0.0f, 1.0f - f * inv_m) * inv_m : 0.0f;
        const float w2 = ok2 ? fmaxf(0.0f, 1.0f - (1.0f - f) * inv_m) * inv_m : 0.0f;
        const int na = min(max(n1, 0), a.nd - 1), nb = min(max(n1 + 1, 0), a.nd - 1);
        const float *base = a.src + (size_t)ang * a.nd * a.nz + z0;
        lerp_accumulate<ZT>(acc, base + (size_t)na * a.nz, base + (size_t)nb * a.nz, w1, w2);
    }
    float *out = a.dst + ((size_t)y * N + x) * a.nz + z0;
#pragma unroll
    for (int g = 0; g < ZT / 4; ++g)
        *reinterpret_cast<float4 *>(out + 4 * g) =
            make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
}

// LDS-staged adjoint: the structure of the parallel back-projector's tiled
// kernel (16 x 16 voxel tile, ZT slices, KC angles staged per chunk, the
// XCD-striped tile order) with the Joseph kernel K as weights.  For each
// angle the tile stages the SPAN detector bins from floor(u) of its lowest
// corner; bins outside [0, n_det) are staged as zeros, which is where the
// direct kernel's clamped loads have weight zero.
constexpr int kJSpan = 26;

template <int ZT, int KC>
__global__ __launch_bounds__(256) void joseph_adjoint_lds_kernel(JosephArgs a, int ntx, int nty, int ntz)
{
    constexpr int T = 16, SPAN = kJSpan, G = ZT / 4;
    __shared__ float sm[KC][SPAN][ZT + 4];
    __shared__ int sb0[KC];
    // block -> tile: XCD c (= block % 8) takes the c-th eighth of the tiles in
    // slice-chunk, 4-column stripe, row order
    const int n = ntx * nty * ntz;
    const int per = (n + 7) >> 3;
    const int b = (int)blockIdx.x;
    const int tidx = (b & 7) * per + (b >> 3);
    if (tidx >= n) return;
    const int plane = ntx * nty;
    const int bz = tidx / plane;
    const int r = tidx - bz * plane;
    const int sw = 4;
    const int s = r / (sw * nty);
    const int q = r - s * sw * nty;
    const int w = min(sw, ntx - sw * s);
    const int by = q / w, bx = sw * s + (q - by * w);

    const int N = a.nx;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = bx * T, y0 = by * T;
    const int x = x0 + tx, y = y0 + ty;
    const int z0 = bz * ZT;
    const bool in_vol = x < N && y < N;
    float acc[ZT];
#pragma unroll
    for (int z = 0; z < ZT; ++z) acc[z] = 0.0f;
    const float xf = (float)(x + 1), yf = (float)(y + 1);
    for (int kb = 0; kb < a.na; kb += KC) {
        const int kc = min(KC, a.na - kb);
        if (threadIdx.x < kc) {
            const float4 t = a.tab[kb + threadIdx.x];
            const float ya = (float)(y0 + 1), yb = (float)(y0 + T);
            const float xa = (float)(x0 + 1), xb = (float)(x0 + T);
            const float la = t.x + t.y * ya, lb = t.x + t.y * yb;
            const float umin = fminf(fminf(la - t.z * xa, la - t.z * xb), fminf(lb - t.z * xa, lb - t.z * xb));
            sb0[threadIdx.x] = (int)floorf(umin) - 1;   // one bin of rounding margin
        }
        __syncthreads();
        for (int e = threadIdx.x; e < kc * SPAN * G; e += 256) {
            const int kk = e / (SPAN * G);
            const int rr = e - kk * SPAN * G;
            const int bb = rr / G, g = rr - bb * G;
            const int col = sb0[kk] + bb;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (col >= 0 && col < a.nd)
                v = *reinterpret_cast<const float4 *>(a.src + ((size_t)(kb + kk) * a.nd + col) * a.nz + z0 + 4 * g);
            *reinterpret_cast<float4 *>(&sm[kk][bb][4 * g]) = v;
        }
        __syncthreads();
        if (in_vol) {
            for (int kk = 0; kk < kc; ++kk) {
                const float4 t = tab_entry(a.tab, kb + kk);
                const float u = (t.x + t.y * yf) - t.z * xf;
                const float fl = floorf(u);
                const int n1 = (int)fl;
                const float f = u - fl;
                const float inv_m = t.w;
                const float w1 = fmaxf(0.0f, 1.0f - f * inv_m) * inv_m;
                const float w2 = fmaxf(0.0f, 1.0f - (1.0f - f) * inv_m) * inv_m;
                const int loc = min(max(n1 - sb0[kk], 0), SPAN - 2);
                const float *A = &sm[kk][loc][0];
                const float *B = A + (ZT + 4);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const float4 va = *reinterpret_cast<const float4 *>(A + 4 * g);
                    const float4 vb = *reinterpret_cast<const float4 *>(B + 4 * g);
                    acc[4 * g + 0] = __builtin_fmaf(w2, vb.x, __builtin_fmaf(w1, va.x, acc[4 * g + 0]));
                    acc[4 * g + 1] = __builtin_fmaf(w2, vb.y, __builtin_fmaf(w1, va.y, acc[4 * g + 1]));
                    acc[4 * g + 2] = __builtin_fmaf(w2, vb.z, __builtin_fmaf(w1, va.z, acc[4 * g + 2]));
                    acc[4 * g + 3] = __builtin_fmaf(w2, vb.w, __builtin_fmaf(w1, va.w, acc[4 * g + 3]));
                }
            }
        }
        __syncthreads();
    }
    if (!in_vol) return;
    float *out = a.dst + ((size_t)y * N + x) * a.nz + z0;
#pragma unroll
    for (int g = 0; g < G; ++g)
        *reinterpret_cast<float4 *>(out + 4 * g) =
            make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
}

std::vector<float> joseph_tables(const mr_joseph_geom &g, const float *angles, size_t n_ang)
{
    std::vector<float> tab(4 * n_ang);
    const double dirw = g.direction == 0 ? -1.0 : 1.0;
    const double cx = (double)(g.nx / 2);
    for (size_t i = 0; i < n_ang; ++i) {
        const double phi = dirw * (double)angles[i] * 3.14159265358979323846 / 180.0;
        const double s = std::sin(phi), c = std::cos(phi);
        tab[4 * i + 0] = (float)(cx * (s - c) + (double)g.center);
        tab[4 * i + 1] = (float)c;
        tab[4 * i + 2] = (float)s;
        tab[4 * i + 3] = (float)(1.0 / std::max(std::fabs(s), std::fabs(c)));
    }
    return tab;
}

// Angle chunks of the LDS forward: angles of one stepping class in order of
// direction, cut so that on every line the crossings of any block's rays
// (kFwdNB bins x the chunk's angles) span at most kFwdW - 3 voxels (one voxel
// of margin for float rounding; the interpolation needs p0 and p0 + 1).  The
// span of affine functions of the line index is largest at the first or
// last line, so those two are checked.
JosephPacked joseph_pack(const mr_joseph_geom &g, const float *angles, size_t n_ang)
{
    JosephPacked pk;
    pk.tab = joseph_tables(g, angles, n_ang);
    const int N = (int)g.nx, nd = (int)g.n_det;
    const int nbt = (nd + kFwdNB - 1) / kFwdNB;
    for (int cls = 0; cls < 2; ++cls) {
        std::vector<std::pair<double, int>> idx;
        for (size_t i = 0; i < n_ang; ++i) {
            const float c = pk.tab[4 * i + 1], s = pk.tab[4 * i + 2];
            const int k = std::fabs(s) >= std::fabs(c) ? 0 : 1;
            if (k != cls) continue;
            double dir = std::atan2((double)s, (double)c);
            if (dir < 0) dir += 2.0 * 3.14159265358979323846;
            idx.push_back({dir, (int)i});
        }
        std::sort(idx.begin(), idx.end());
        // per bin tile and end line: min / max crossing over the chunk
        std::vector<double> mn(2 * nbt), mx(2 * nbt), tmn(2 * nbt), tmx(2 * nbt);
        int first = (int)pk.order.size(), count = 0;
        auto flush = [&]() {
            if (count > 0) pk.chunks.push_back(make_int4(first, count, cls, 0));
            first = (int)pk.order.size();
            count = 0;
        };
        for (const auto &e : idx) {
            const int i = e.second;
            const double su = pk.tab[4 * i], c = pk.tab[4 * i + 1], s = pk.tab[4 * i + 2];
            const double a0 = cls == 0 ? su : -su, b1 = cls == 0 ? c : s, inv = cls == 0 ? 1.0 / s : 1.0 / c;
            double worst = 0.0;
            for (int bt = 0; bt < nbt; ++bt) {
                const int na = bt * kFwdNB, nb = std::min(na + kFwdNB - 1, nd - 1);
                for (int end = 0; end < 2; ++end) {
                    const double L1 = end == 0 ? 1.0 : (double)N;
                    double lo = 1e300, hi = -1e300;
                    for (int nn : {na, nb}) {
                        const double cn = cls == 0 ? -(double)nn : (double)nn;
                        const double pf = ((a0 + b1 * L1) + cn) * inv - 1.0;
                        lo = std::min(lo, pf);
                        hi = std::max(hi, pf);
                    }
                    const int k = 2 * bt + end;
                    tmn[k] = count ? std::min(mn[k], lo) : lo;
                    tmx[k] = count ? std::max(mx[k], hi) : hi;
                    worst = std::max(worst, tmx[k] - tmn[k]);
                }
            }
            if (count > 0 && (count == kFwdKA || std::ceil(worst) + 1.0 > (double)(kFwdW - 2))) {
                flush();
                // restart the extents with this angle alone
                for (int bt = 0; bt < nbt; ++bt) {
                    const int na = bt * kFwdNB, nb = std::min(na + kFwdNB - 1, nd - 1);
                    for (int end = 0; end < 2; ++end) {
                        const double L1 = end == 0 ? 1.0 : (double)N;
                        double lo = 1e300, hi = -1e300;
                        for (int nn : {na, nb}) {
                            const double cn = cls == 0 ? -(double)nn : (double)nn;
                            const double pf = ((a0 + b1 * L1) + cn) * inv - 1.0;
                            lo = std::min(lo, pf);
                            hi = std::max(hi, pf);
                        }
                        tmn[2 * bt + end] = lo;
                        tmx[2 * bt + end] = hi;
                    }
                }
            }
            mn = tmn;
            mx = tmx;
            pk.order.push_back(i);
            ++count;
        }
        flush();
    }
    return pk;
}

size_t JosephPacked::bytes() const
{
    return sizeof(float) * tab.size() + sizeof(int4) * chunks.size() + sizeof(int) * order.size();
}

// One buffer: tables, then chunks, then the angle order.
int joseph_upload(const JosephPacked &pk, void *d_base, hipStream_t s, JosephDevTables &dt)
{
    char *b = static_cast<char *>(d_base);
    const size_t o1 = sizeof(float) * pk.tab.size(), o2 = o1 + sizeof(int4) * pk.chunks.size();
    MR_HIP(hipMemcpyAsync(b, pk.tab.data(), o1, hipMemcpyHostToDevice, s));
    if (!pk.chunks.empty())
        MR_HIP(hipMemcpyAsync(b + o1, pk.chunks.data(), o2 - o1, hipMemcpyHostToDevice, s));
    if (!pk.order.empty())
        MR_HIP(hipMemcpyAsync(b + o2, pk.order.data(), sizeof(int) * pk.order.size(), hipMemcpyHostToDevice, s));
    dt.tab = reinterpret_cast<const float4 *>(b);
    dt.chunks = reinterpret_cast<const int4 *>(b + o1);
    dt.order = reinterpret_cast<const int *>(b + o2);
    dt.n_chunks = (int)pk.chunks.size();
    return MR_OK;
}

int joseph_check(const mr_joseph_geom *g, size_t n_ang)
{
    MR_REQUIRE(g != nullptr, "geometry is NULL");
    MR_REQUIRE(g->nx >= 2 && g->n_det >= 2 && g->nz >= 4, "sizes too small");
    MR_REQUIRE(g->nz % 4 == 0, "nz must be a multiple of 4");
    MR_REQUIRE(n_ang >= 1 && n_ang <= 65535 * 4, "angle count out of range");
    MR_REQUIRE(g->direction == 0 || g->direction == 1, "direction must be 0 (CW) or 1 (CCW)");
    return MR_OK;
}

// Enqueues A (forward) or A^T on s with a resident angle table; no host sync.
int joseph_launch(const mr_joseph_geom *g, const float *src, float *dst, const JosephDevTables &dt, size_t n_ang,
                  bool forward, hipStream_t s)
{
    mr::JosephArgs a{src, dst, dt.tab, (int)g->nx, (int)g->nz, (int)g->n_det, (int)n_ang};
    constexpr int ZT = 4;
    // ZT = 16 when the slice count allows it
    if (forward) {
        mr::ProfScope prof("joseph_forward", s);
        if (g->nz % 16 == 0 && dt.n_chunks > 0) {
            // 32 slices per block when the slice count allows (measured at
            // 1024^2 x 64 x 900: 7.7 ms against 8.4 with 16)
            const int zt = g->nz % 32 == 0 ? 32 : 16;
            const int nbt = (int)((g->n_det + kFwdNB - 1) / kFwdNB), nzs = (int)(g->nz / zt);
            const long total = (long)dt.n_chunks * nbt * nzs;
            MR_REQUIRE(total < (1L << 31) - 8, "Joseph forward grid too large");
            const dim3 grid((unsigned)(8 * ((total + 7) / 8)));
            if (g->nx <= 4096) {
                // one barrier per line (spans precomputed in LDS, 8 bytes per line)
                const size_t dyn = sizeof(int) * 2 * g->nx;
                if (zt == 32)
                    hipLaunchKernelGGL((mr::joseph_forward_lds2_kernel<32>), grid, dim3(256), dyn, s, a, dt.chunks,
                                       dt.order, dt.n_chunks, nbt, nzs);
                else
                    hipLaunchKernelGGL((mr::joseph_forward_lds2_kernel<16>), grid, dim3(256), dyn, s, a, dt.chunks,
                                       dt.order, dt.n_chunks, nbt, nzs);
            } else if (zt == 32) {
                // wider volumes: the span table would not fit beside the staging
                // buffers; the per-line form gives the same bits
                hipLaunchKernelGGL((mr::joseph_forward_lds_kernel<32, kFwdR>), grid, dim3(256), 0, s, a, dt.chunks,
                                   dt.order, dt.n_chunks, nbt, nzs);
            } else {
                hipLaunchKernelGGL((mr::joseph_forward_lds_kernel<16, kFwdR>), grid, dim3(256), 0, s, a, dt.chunks,
                                   dt.order, dt.n_chunks, nbt, nzs);
            }
        } else {
            // slice counts that are not a multiple of 16: direct ray-driven gathers
            dim3 grid((unsigned)((g->n_det + 63) / 64), (unsigned)((n_ang + 3) / 4), (unsigned)(g->nz / ZT));
            hipLaunchKernelGGL(mr::joseph_forward_kernel<ZT>, grid, dim3(256), 0, s, a);
        }
    } else {
        mr::ProfScope prof("joseph_adjoint", s);
        const int ntx = (int)((g->nx + 15) / 16);
        if (g->nz % 32 == 0) {
            const long nt = (long)ntx * ntx * (long)(g->nz / 32);
            hipLaunchKernelGGL((mr::joseph_adjoint_lds_kernel<32, 4>), dim3((unsigned)(8 * ((nt + 7) / 8))), dim3(256),
                               0, s, a, ntx, ntx, (int)(g->nz / 32));
        } else if (g->nz % 16 == 0) {
            const long nt = (long)ntx * ntx * (long)(g->nz / 16);
            hipLaunchKernelGGL((mr::joseph_adjoint_lds_kernel<16, 8>), dim3((unsigned)(8 * ((nt + 7) / 8))), dim3(256),
                               0, s, a, ntx, ntx, (int)(g->nz / 16));
        } else {
            // slice counts that are not a multiple of 16: voxel-driven gathers
            dim3 grid((unsigned)((g->nx + 15) / 16), (unsigned)((g->nx + 15) / 16), (unsigned)(g->nz / ZT));
            hipLaunchKernelGGL(mr::joseph_adjoint_kernel<ZT>, grid, dim3(256), 0, s, a);
        }
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr

namespace {

int joseph_common(const mr_joseph_geom *g, const float *src, const float *angles, size_t n_ang, float *dst,
                  void *stream, bool forward)
{
    mr::clear_error();
    int rc = mr::joseph_check(g, n_ang);
    if (rc != MR_OK) return rc;
    MR_REQUIRE(src != nullptr && dst != nullptr && angles != nullptr, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    const mr::JosephPacked pk = mr::joseph_pack(*g, angles, n_ang);
    void *d_buf = nullptr;
    MR_HIP(hipMallocAsync(&d_buf, pk.bytes(), s));
    mr::JosephDevTables dt;
    rc = mr::joseph_upload(pk, d_buf, s, dt);
    if (rc == MR_OK) rc = mr::joseph_launch(g, src, dst, dt, n_ang, forward, s);
    hipError_t e2 = hipFreeAsync(d_buf, s);
    if (rc != MR_OK) return rc;
    if (e2 != hipSuccess) return mr::hip_fail(e2, "hipFreeAsync", __FILE__, __LINE__);
    // the host table must outlive the asynchronous copy
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
}

}  // namespace

extern "C" int mr_joseph_forward(const mr_joseph_geom *geom, const float *d_volume, const float *angles,
                                 size_t n_ang, float *d_sino, void *stream)
{
    MR_GUARD_BEGIN
    return joseph_common(geom, d_volume, angles, n_ang, d_sino, stream, true);
    MR_GUARD_END
}

extern "C" int mr_joseph_adjoint(const mr_joseph_geom *geom, const float *d_sino, const float *angles,
                                 size_t n_ang, float *d_volume, void *stream)
{
    MR_GUARD_BEGIN
    return joseph_common(geom, d_sino, angles, n_ang, d_volume, stream, false);
    MR_GUARD_END
}
