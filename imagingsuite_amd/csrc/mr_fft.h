// LDS-resident Stockham FFT for one complex line of compile-time length P,
// shared by the parallel-beam and the FDK ramp filters.
//
// The line lives in LDS as float2[P].  Passes are radix-4 (plus one radix-2
// pass first when log2 P is odd); every pass reads its butterfly inputs into
// registers, synchronises, and writes the outputs back in natural (Stockham
// autosort) order, so no bit reversal is needed.  Twiddles come from a
// P-entry table tw[k] = exp(-2 pi i k / P) computed in double on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace mr {

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

template <int P, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2 *a, const float2 *__restrict__ tw)
{
    constexpr int L = ilog2c(P);
    const int t = threadIdx.x;
    int ns = 1;
    int lns = 0;
    if constexpr ((L & 1) != 0) {
        constexpr int B = P / 2;
        constexpr int PER = (B + NT - 1) / NT;
        float2 v0[PER], v1[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = t + q * NT;
            if (j < B) { v0[q] = a[j]; v1[q] = a[j + B]; }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = t + q * NT;
            if (j < B) {
                a[2 * j] = cadd(v0[q], v1[q]);
                a[2 * j + 1] = csub(v0[q], v1[q]);
            }
        }
        __syncthreads();
        ns = 2;
        lns = 1;
    }
    constexpr int B4 = P / 4;
    constexpr int PER4 = (B4 + NT - 1) / NT;
    for (; ns < P; ns <<= 2, lns += 2) {
        float2 v[PER4][4];
        const int tshift = L - 2 - lns;  // log2(P / (4 ns))
#pragma unroll
        for (int q = 0; q < PER4; ++q) {
            const int j = t + q * NT;
            if (j < B4) {
                const int k = j & (ns - 1);
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][r] = a[j + r * B4];
                const int base = k << tshift;
#pragma unroll
                for (int r = 1; r < 4; ++r) {
                    float2 w = tw[base * r];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
                const float2 a0 = cadd(v[q][0], v[q][2]);
                const float2 a1 = csub(v[q][0], v[q][2]);
                const float2 a2 = cadd(v[q][1], v[q][3]);
                const float2 d = csub(v[q][1], v[q][3]);
                const float2 a3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
                v[q][0] = cadd(a0, a2);
                v[q][2] = csub(a0, a2);
                v[q][1] = cadd(a1, a3);
                v[q][3] = csub(a1, a3);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER4; ++q) {
            const int j = t + q * NT;
            if (j < B4) {
                const int k = j & (ns - 1);
                const int d0 = ((j >> lns) << (lns + 2)) + k;
#pragma unroll
                for (int r = 0; r < 4; ++r) a[d0 + r * ns] = v[q][r];
            }
        }
        __syncthreads();
    }
}

}  // namespace mr
