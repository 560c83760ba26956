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

// ---------------------------------------------------------------------------
// Second LDS FFT: radix-8 Stockham passes (one radix-2 or radix-4 pass first
// when log2 P is not a multiple of 3), so P = 8192 takes 5 passes instead of
// 7.  The line is stored with one pad entry every 16 (index lpad(i)), which
// spreads the strided writes of the early passes over all LDS banks.  The
// twiddles come from an LDS copy of the first P/8 entries of the table
// (w = exp(-+2 pi i base / P), base < P/8 in every radix-8 pass); the other
// six powers of a butterfly are products of w (at most three multiplications
// deep).  Length of the padded line: lds_len(P).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
constexpr int lds_len(int P) { return P + P / 16; }

template <bool INV>
__device__ __forceinline__ void dft4(float2 &v0, float2 &v1, float2 &v2, float2 &v3)
{
    const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), d = csub(v1, v3);
    const float2 a3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_r(float2 (&v)[R])
{
    if constexpr (R == 2) {
        const float2 t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else {
        constexpr float c = 0.70710678118654752440f;
        float2 a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            a[r] = cadd(v[r], v[r + 4]);
            b[r] = csub(v[r], v[r + 4]);
        }
        // b_r *= exp(-+2 pi i r / 8)
        b[1] = INV ? make_float2(c * (b[1].x - b[1].y), c * (b[1].x + b[1].y))
                   : make_float2(c * (b[1].x + b[1].y), c * (b[1].y - b[1].x));
        b[2] = INV ? make_float2(-b[2].y, b[2].x) : make_float2(b[2].y, -b[2].x);
        b[3] = INV ? make_float2(-c * (b[3].x + b[3].y), c * (b[3].x - b[3].y))
                   : make_float2(c * (b[3].y - b[3].x), -c * (b[3].x + b[3].y));
        dft4<INV>(a[0], a[1], a[2], a[3]);
        dft4<INV>(b[0], b[1], b[2], b[3]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            v[2 * m] = a[m];
            v[2 * m + 1] = b[m];
        }
    }
}

template <int P, int NT, int R, bool INV>
__device__ __forceinline__ void fft_pass(float2 *a, const float2 *__restrict__ twl, int ns, int lns)
{
    constexpr int B = P / R;
    constexpr int PER = (B + NT - 1) / NT;
    constexpr int LR = ilog2c(R);
    constexpr int L = ilog2c(P);
    const int t = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * NT;
        if (j < B) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = a[lpad(j + r * B)];
            if (R == 8 && ns > 1) {
                const int k = j & (ns - 1);
                float2 w1 = twl[k << (L - 3 - lns)];
                if (INV) w1.y = -w1.y;
                const float2 w2 = cmul(w1, w1), w3 = cmul(w1, w2), w4 = cmul(w2, w2);
                const float2 w5 = cmul(w1, w4), w6 = cmul(w3, w3), w7 = cmul(w3, w4);
                const float2 w[8] = {make_float2(1.f, 0.f), w1, w2, w3, w4, w5, w6, w7};
#pragma unroll
                for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], w[r]);
            }
            dft_r<R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * NT;
        if (j < B) {
            const int k = j & (ns - 1);
            const int d0 = ((j >> lns) << (lns + LR)) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) a[lpad(d0 + r * ns)] = v[q][r];
        }
    }
    __syncthreads();
}

template <int P, int NT, bool INV>
__device__ __forceinline__ void fft_lds2(float2 *a, const float2 *__restrict__ twl)
{
    constexpr int L = ilog2c(P);
    int ns = 1, lns = 0;
    if constexpr (L % 3 == 1) {
        fft_pass<P, NT, 2, INV>(a, twl, 1, 0);
        ns = 2;
        lns = 1;
    } else if constexpr (L % 3 == 2) {
        fft_pass<P, NT, 4, INV>(a, twl, 1, 0);
        ns = 4;
        lns = 2;
    }
    for (; ns < P; ns <<= 3, lns += 3) fft_pass<P, NT, 8, INV>(a, twl, ns, lns);
}

// The same passes with the first pass's inputs and the last pass's outputs
// taken from / handed to functors instead of the LDS line, so that a filter
// can build its padded input in the first pass, multiply by its spectrum in
// the last forward pass and crop/scale in the last inverse pass without an
// LDS round trip (and its barriers) of its own.  load(i) returns input i;
// store(i, v) receives output i.  LOAD_F / STORE_F select the functors for
// the first / last pass; STORE_LDS says that the store functor writes the
// LDS line (then the pass ends with a barrier, as an LDS pass does).
template <int P, int NT, int R, bool INV, bool LOAD_F, bool STORE_F, bool STORE_LDS, class Load, class Store>
__device__ __forceinline__ void fft_pass_io(float2 *a, const float2 *__restrict__ twl, int ns, int lns, Load load,
                                            Store store)
{
    constexpr int B = P / R;
    constexpr int PER = (B + NT - 1) / NT;
    constexpr int LR = ilog2c(R);
    constexpr int L = ilog2c(P);
    const int t = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * NT;
        if (j < B) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = LOAD_F ? load(j + r * B) : a[lpad(j + r * B)];
            if (R == 8 && ns > 1) {
                const int k = j & (ns - 1);
                float2 w1 = twl[k << (L - 3 - lns)];
                if (INV) w1.y = -w1.y;
                const float2 w2 = cmul(w1, w1), w3 = cmul(w1, w2), w4 = cmul(w2, w2);
                const float2 w5 = cmul(w1, w4), w6 = cmul(w3, w3), w7 = cmul(w3, w4);
                const float2 w[8] = {make_float2(1.f, 0.f), w1, w2, w3, w4, w5, w6, w7};
#pragma unroll
                for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], w[r]);
            }
            dft_r<R, INV>(v[q]);
        }
    }
    // in place: every read of the line must be done before it is overwritten
    if (!LOAD_F && !(STORE_F && !STORE_LDS)) __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * NT;
        if (j < B) {
            const int k = j & (ns - 1);
            const int d0 = ((j >> lns) << (lns + LR)) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (STORE_F) store(d0 + r * ns, v[q][r]);
                else a[lpad(d0 + r * ns)] = v[q][r];
            }
        }
    }
    if (!STORE_F || STORE_LDS) __syncthreads();
}

// FFT of the line whose first pass reads load(i) (the line in LDS is
// written by that pass) and whose last pass hands every output to store(i, v)
// (STORE_LDS: store writes the LDS line, which the caller reads afterwards).
template <int P, int NT, bool INV, bool LOAD_F, bool STORE_F, bool STORE_LDS, class Load, class Store>
__device__ __forceinline__ void fft_lds2_io(float2 *a, const float2 *__restrict__ twl, Load load, Store store)
{
    constexpr int L = ilog2c(P);
    constexpr int FIRST = L % 3 == 1 ? 2 : (L % 3 == 2 ? 4 : 8);   // radix of the first pass
    constexpr int NPASS = (L % 3 == 0 ? 0 : 1) + L / 3;
    auto nothing_load = [](int) { return make_float2(0.f, 0.f); };
    auto nothing_store = [](int, float2) {};
    if constexpr (NPASS == 1) {
        fft_pass_io<P, NT, FIRST, INV, LOAD_F, STORE_F, STORE_LDS>(a, twl, 1, 0, load, store);
        return;
    }
    fft_pass_io<P, NT, FIRST, INV, LOAD_F, false, false>(a, twl, 1, 0, load, nothing_store);
    int ns = FIRST, lns = ilog2c(FIRST);
#pragma unroll 1
    for (int pass = 1; pass < NPASS - 1; ++pass, ns <<= 3, lns += 3)
        fft_pass_io<P, NT, 8, INV, false, false, false>(a, twl, ns, lns, nothing_load, nothing_store);
    fft_pass_io<P, NT, 8, INV, false, STORE_F, STORE_LDS>(a, twl, ns, lns, nothing_load, store);
}

}  // namespace mr
