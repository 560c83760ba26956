// Ramp filters on gfx950: ProjectionFilterSingle (parallel beam) and
// FDKbp_single::ramp_filter_tuned (cone beam).
//
// One workgroup filters two detector rows at once: row a goes into the real
// part and row b into the imaginary part of one complex line of length P held
// in LDS.  Both filters multiply the spectrum by a real, even sequence, so
// IFFT(Hs * FFT(a + i b)) = filt(a) + i filt(b) exactly and no split/merge
// step is needed.  Padding, tapering and scaling are fused into the load and
// store of the line; the rows are filtered in place in HBM.
#include "mr_common.h"
#include "mr_fft.h"
#include "mr_host.h"
#include "mr_filter.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace mr {



// ProjectionFilterSingle::Pad (ProjectionFilter.cpp:346-361) evaluated at
// position i < P of the P+16 long line.
__device__ __forceinline__ float pb_pad(const float *row, int n, int i, int ins, int P,
                                        const float *__restrict__ taper, float first, float last)
{
    if (i < ins) return first * taper[i];
    const int d = i - ins;
    if (d < n) return row[d];
    const int t = P + 15 - i;
    return (t < ins) ? last * taper[t] : 0.0f;
}

// ramp_filter_tuned zero padding with the 5-column edge copies
// (fdkbp_single.cpp:855-897); the row itself is pre-multiplied by the
// projection weight (FdkReconBase::Process, fdkreconbase.cpp:243-249).
// With a.cosw set (an extension without reference counterpart) each sample
// is also multiplied by the Feldkamp cosine weight SDD / sqrt(SDD^2 + u^2 + v^2),
// u = (d - cw_c0) pix_u and v = (row - cw_r0) pix_v in mm from the piercing
// point; vv2 = (v / SDD)^2 of the row.  The 5-row edge copies run after the
// filter, so the copied rows carry the weight of the row they copy.
__device__ __forceinline__ float fdk_pad(const float *row, int n, int c, int ins, int P, float w,
                                         const RowFilterArgs &a, float vv2)
{
    const int cc = c < 5 ? 5 : (c >= P - 5 ? P - 6 : c);
    const int d = cc - ins;
    if (d < 0 || d >= n) return 0.0f;
    const float x = row[d] * w;
    if (!a.cosw) return x;
    const float uu = ((float)d - a.cw_c0) * a.cw_u;
    return x * (1.0f / sqrtf(fmaf(uu, uu, 1.0f + vv2)));
}

__device__ __forceinline__ float fdk_row_vv2(const RowFilterArgs &a, long r)
{
    const float vv = ((float)(r % a.rows_per_proj) - a.cw_r0) * a.cw_v;
    return vv * vv;
}

// The filter on the radix-8 padded FFT (fft_lds2): dynamic LDS holds the
// padded line (lds_len(P) float2) followed by the P/8 twiddles it uses.
template <int P, int NT>
__global__ __launch_bounds__(NT) void row_filter_kernel2(RowFilterArgs a)
{
    extern __shared__ float2 line[];
    float2 *twl = line + lds_len(P);
    const long r0 = 2 * (long)blockIdx.x;
    const bool has1 = r0 + 1 < a.n_rows;
    float *row0 = a.rows + r0 * a.n;
    float *row1 = has1 ? row0 + a.n : row0;
    const int t = threadIdx.x;
    for (int i = t; i < P / 8; i += NT) twl[i] = a.tw[i];
    float *o0 = a.out ? a.out + r0 * a.n : row0;
    float *o1 = o0 + a.n;
    if (a.mode != 2) {
        // The padded line is built in the first pass's loads, the spectrum
        // multiplied in the last forward pass and the kept samples cropped
        // and scaled in the last inverse pass: no LDS round trip of their
        // own.  Same values and operations as staging the line first.
        float f0 = 0.f, l0 = 0.f, f1 = 0.f, l1 = 0.f, w0 = 1.f, w1 = 1.f, v0 = 0.f, v1 = 0.f;
        float sa0 = 0.f, sa1 = 0.f, sb0 = 0.f, sb1 = 0.f;   // mode 4: the rows' dot products with the responses
        // mode 3 with norm: row r is row r % rows_per_proj of projection r / rows_per_proj
        const float *nf = a.norm_flat_log, *nd = a.norm_dark;
        size_t pix0 = 0, pix1 = 0;
        float dl0 = 0.f, dl1 = 0.f;
        if (a.mode == 3 && a.norm) {
            const long rr1 = has1 ? r0 + 1 : r0;
            pix0 = (size_t)(r0 % a.rows_per_proj) * a.n;
            pix1 = (size_t)(rr1 % a.rows_per_proj) * a.n;
            dl0 = a.norm_dl ? a.norm_dl[r0 / a.rows_per_proj] : 0.0f;
            dl1 = a.norm_dl ? a.norm_dl[rr1 / a.rows_per_proj] : 0.0f;
        }
        auto raw0 = [&](int i) { return a.norm ? lognorm_px(row0[i], nf, nd, pix0 + i, dl0) : row0[i]; };
        auto raw1 = [&](int i) { return a.norm ? lognorm_px(row1[i], nf, nd, pix1 + i, dl1) : row1[i]; };
        if (a.mode == 3) {
            f0 = raw0(0); l0 = raw0(a.n - 1);
            f1 = raw1(0); l1 = raw1(a.n - 1);
        } else if (a.mode == 0) {
            f0 = row0[0]; l0 = row0[a.n - 1];
            f1 = row1[0]; l1 = row1[a.n - 1];
        } else {
            w0 = a.pre_w ? a.pre_w[r0 / a.rows_per_proj] : 1.0f;
            w1 = (a.pre_w && has1) ? a.pre_w[(r0 + 1) / a.rows_per_proj] : 1.0f;
            v0 = fdk_row_vv2(a, r0);
            v1 = fdk_row_vv2(a, r0 + 1);
        }
        auto load = [&](int i) -> float2 {
            if (a.mode == 4) {
                if (i >= a.n) return make_float2(0.0f, 0.0f);
                const float x0 = row0[i], x1 = has1 ? row1[i] : 0.0f;
                const float ra = a.resp_first[i], rb = a.resp_last[i];
                sa0 += ra * x0;
                sa1 += ra * x1;
                sb0 += rb * x0;
                sb1 += rb * x1;
                return make_float2(x0, x1);
            }
            if (a.mode == 3) {
                if (i >= a.n) return make_float2(0.0f, 0.0f);
                const float s = (float)i * a.ramp_inv, t1 = 1.0f - s;
                return make_float2(raw0(i) - (f0 * t1 + l0 * s), has1 ? raw1(i) - (f1 * t1 + l1 * s) : 0.0f);
            }
            if (a.mode == 0)
                return make_float2(pb_pad(row0, a.n, i, a.insert, P, a.taper, f0, l0),
                                   has1 ? pb_pad(row1, a.n, i, a.insert, P, a.taper, f1, l1) : 0.0f);
            return make_float2(fdk_pad(row0, a.n, i, a.insert, P, w0, a, v0),
                               has1 ? fdk_pad(row1, a.n, i, a.insert, P, w1, a, v1) : 0.0f);
        };
        auto spectrum = [&](int i, float2 v) {
            const float h = a.hs[i];
            line[lpad(i)] = make_float2(v.x * h, v.y * h);
        };
        __syncthreads();   // the twiddles
        fft_lds2_io<P, NT, false, true, true, true>(line, twl, load, spectrum);
        __shared__ float red4[NT / 64 + 1][4];
        float A0 = 0.f, A1 = 0.f, B0 = 0.f, B1 = 0.f;
        if (a.mode == 4) {
            // block sums of the dot products, read after the barrier below
            for (int off = 32; off > 0; off >>= 1) {
                sa0 += __shfl_xor(sa0, off);
                sa1 += __shfl_xor(sa1, off);
                sb0 += __shfl_xor(sb0, off);
                sb1 += __shfl_xor(sb1, off);
            }
            if ((t & 63) == 0) {
                red4[t >> 6][0] = sa0;
                red4[t >> 6][1] = sa1;
                red4[t >> 6][2] = sb0;
                red4[t >> 6][3] = sb1;
            }
            __syncthreads();
            for (int wv = 0; wv < (NT + 63) / 64; ++wv) {
                A0 += red4[wv][0];
                A1 += red4[wv][1];
                B0 += red4[wv][2];
                B1 += red4[wv][3];
            }
        }
        auto none = [](int) { return make_float2(0.f, 0.f); };
        auto crop = [&](int i, float2 v) {
            if (a.mode == 4) {
                if (i >= a.n) return;
                float y0 = v.x, y1 = v.y;
                if (i == 0) { y0 += A0; y1 += A1; }
                if (i == a.n - 1) { y0 += B0; y1 += B1; }
                o0[i] = y0;
                if (has1) o1[i] = y1;
                return;
            }
            if (a.mode == 3) {
                if (i >= a.n) return;
                const float rf = a.resp_first[i], rl = a.resp_last[i];
                o0[i] = (v.x + f0 * rf) + l0 * rl;
                if (has1) o1[i] = (v.y + f1 * rf) + l1 * rl;
                return;
            }
            const int d = i - a.insert;
            if (d < 0 || d >= a.n) return;
            if (a.mode == 0) {
                o0[d] = v.x * a.post0;
                if (has1) o1[d] = v.y * a.post0;
            } else {
                o0[d] = ((v.x * a.post0) * a.post1) * a.post2;
                if (has1) o1[d] = ((v.y * a.post0) * a.post1) * a.post2;
            }
        };
        fft_lds2_io<P, NT, true, false, true, false>(line, twl, none, crop);
        return;
    }
    if (a.mode == 0) {
        const float f0 = row0[0], l0 = row0[a.n - 1];
        const float f1 = row1[0], l1 = row1[a.n - 1];
        for (int i = t; i < P; i += NT) {
            const float x0 = pb_pad(row0, a.n, i, a.insert, P, a.taper, f0, l0);
            const float x1 = has1 ? pb_pad(row1, a.n, i, a.insert, P, a.taper, f1, l1) : 0.0f;
            line[lpad(i)] = make_float2(x0, x1);
        }
    } else if (a.mode == 1) {
        const float w0 = a.pre_w ? a.pre_w[r0 / a.rows_per_proj] : 1.0f;
        const float w1 = (a.pre_w && has1) ? a.pre_w[(r0 + 1) / a.rows_per_proj] : 1.0f;
        const float v0 = fdk_row_vv2(a, r0), v1 = fdk_row_vv2(a, r0 + 1);
        for (int i = t; i < P; i += NT) {
            const float x0 = fdk_pad(row0, a.n, i, a.insert, P, w0, a, v0);
            const float x1 = has1 ? fdk_pad(row1, a.n, i, a.insert, P, w1, a, v1) : 0.0f;
            line[lpad(i)] = make_float2(x0, x1);
        }
    } else {
        // mode 2, adjoint of mode 0: the scaled rows embedded at nInsert, zeros elsewhere
        for (int i = t; i < P; i += NT) {
            const int d = i - a.insert;
            const bool in = d >= 0 && d < a.n;
            const float x0 = in ? row0[d] * a.post0 : 0.0f;
            const float x1 = (in && has1) ? row1[d] * a.post0 : 0.0f;
            line[lpad(i)] = make_float2(x0, x1);
        }
    }
    __syncthreads();
    fft_lds2<P, NT, false>(line, twl);
    for (int i = t; i < P; i += NT) {
        const float h = a.hs[i];
        const float2 v = line[lpad(i)];
        line[lpad(i)] = make_float2(v.x * h, v.y * h);
    }
    __syncthreads();
    fft_lds2<P, NT, true>(line, twl);
    if (a.mode == 0) {
        for (int i = t; i < a.n; i += NT) {
            const float2 v = line[lpad(i + a.insert)];
            o0[i] = v.x * a.post0;
            if (has1) o1[i] = v.y * a.post0;
        }
    } else if (a.mode == 2) {
        // adjoint of Pad: the tapers fold back onto the first and last sample
        // (left taper at 0..nInsert-1, right taper at P+15-j, inside the
        // FFT window only for j >= 16)
        __shared__ float red[NT / 64 + 1][4];
        float sl0 = 0.0f, sl1 = 0.0f, sr0 = 0.0f, sr1 = 0.0f;
        for (int j = t; j < a.insert; j += NT) {
            const float w = a.taper[j];
            const float2 l = line[lpad(j)];
            sl0 += w * l.x;
            sl1 += w * l.y;
            const int jj = P + 15 - j;
            if (jj < P) {
                const float2 r = line[lpad(jj)];
                sr0 += w * r.x;
                sr1 += w * r.y;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            sl0 += __shfl_xor(sl0, off);
            sl1 += __shfl_xor(sl1, off);
            sr0 += __shfl_xor(sr0, off);
            sr1 += __shfl_xor(sr1, off);
        }
        if ((t & 63) == 0) {
            red[t >> 6][0] = sl0;
            red[t >> 6][1] = sl1;
            red[t >> 6][2] = sr0;
            red[t >> 6][3] = sr1;
        }
        __syncthreads();
        float L0 = 0.0f, L1 = 0.0f, R0 = 0.0f, R1 = 0.0f;
        for (int wv = 0; wv < (NT + 63) / 64; ++wv) {
            L0 += red[wv][0];
            L1 += red[wv][1];
            R0 += red[wv][2];
            R1 += red[wv][3];
        }
        for (int i = t; i < a.n; i += NT) {
            const float2 v = line[lpad(i + a.insert)];
            float y0 = v.x, y1 = v.y;
            if (i == 0) { y0 += L0; y1 += L1; }
            if (i == a.n - 1) { y0 += R0; y1 += R1; }
            o0[i] = y0;
            if (has1) o1[i] = y1;
        }
    } else {
        for (int i = t; i < a.n; i += NT) {
            const float2 v = line[lpad(i + a.insert)];
            o0[i] = ((v.x * a.post0) * a.post1) * a.post2;
            if (has1) o1[i] = ((v.y * a.post0) * a.post1) * a.post2;
        }
    }
}

// FDK margin rows: rows 0..4 take the filtered row 5, rows h-5..h-1 the
// filtered row h-6 (fdkbp_single.cpp:880-886).  Filtering is row-local, so
// copying the filtered rows equals filtering the copied raw rows.
__global__ void fdk_margin_rows_kernel(float *img, int n, int h, long n_proj)
{
    const long p = blockIdx.y;
    const int rr = blockIdx.z;  // 0..9
    const int dst = rr < 5 ? rr : h - 10 + rr;
    const int src = rr < 5 ? 5 : h - 6;
    float *base = img + p * (long)n * h;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        base[(long)dst * n + c] = base[(long)src * n + c];
}

template <int P>
static int launch_row_filter(const RowFilterArgs &a, hipStream_t s)
{
    // one radix-8 butterfly per thread and pass: P / 8 threads (at most
    // 1024); measured for the 384 000 rows of 2048 at P = 8192: 6.75 ms with
    // the padding, spectrum and crop fused into the first and last passes,
    // 7.8-8.1 ms without, 9.3-9.8 ms with P / 16 threads, 19.75 ms with a
    // radix-4 FFT; the line split into 4 or 8 wave-local FFTs of length P / 4
    // or P / 8 (two block barriers per row pair instead of two per pass)
    // took 11.0 / 8.4 ms (2 / 4 waves per SIMD); radix-16 passes with P / 16
    // threads 6.52 against 6.49 ms
    constexpr int NT2 = (P / 8) < 64 ? 64 : ((P / 8) > 1024 ? 1024 : (P / 8));
    const long pairs = (a.n_rows + 1) / 2;
    if (pairs == 0) return MR_OK;
    if (pairs > 0x7fffffffL) {
        set_error("too many rows for one launch");
        return MR_ERR_INVALID;
    }
    ProfScope prof((a.mode == 0 || a.mode == 3) ? "row_filter_pb"
                                                : (a.mode == 1 ? "row_filter_fdk" : "row_filter_adjoint"),
                   s);
    MR_REQUIRE((a.mode != 3 && a.mode != 4) || (a.resp_first && a.resp_last), "short filter without responses");
    const size_t lds = sizeof(float2) * (lds_len(P) + P / 8);
    if (lds > 65536)
        MR_HIP(hipFuncSetAttribute((const void *)row_filter_kernel2<P, NT2>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((row_filter_kernel2<P, NT2>), dim3((unsigned)pairs), dim3(NT2), lds, s, a);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int run_row_filter(const RowFilterArgs &a, int P, hipStream_t s)
{
    switch (P) {
    case 64: return launch_row_filter<64>(a, s);
    case 128: return launch_row_filter<128>(a, s);
    case 256: return launch_row_filter<256>(a, s);
    case 512: return launch_row_filter<512>(a, s);
    case 1024: return launch_row_filter<1024>(a, s);
    case 2048: return launch_row_filter<2048>(a, s);
    case 4096: return launch_row_filter<4096>(a, s);
    case 8192: return launch_row_filter<8192>(a, s);
    case 16384: return launch_row_filter<16384>(a, s);
    default:
        set_error("FFT length " + std::to_string(P) + " outside the supported 64..16384");
        return MR_ERR_UNSUPPORTED;
    }
}

int run_fdk_margin_rows(float *img, int n, int h, long n_proj, hipStream_t s)
{
    if (n_proj == 0) return MR_OK;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n_proj, 10);
    ProfScope prof("fdk_margin_rows", s);
    hipLaunchKernelGGL(fdk_margin_rows_kernel, grid, dim3(256), 0, s, img, n, h, n_proj);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// P twiddles exp(-2 pi i k / P) in double, rounded to float.
std::vector<float> make_twiddles(size_t P)
{
    std::vector<float> tw(2 * P);
    const double pi = 3.1415926535897932384626433832795;
    for (size_t k = 0; k < P; ++k) {
        const double ang = -2.0 * pi * (double)k / (double)P;
        tw[2 * k] = (float)std::cos(ang);
        tw[2 * k + 1] = (float)std::sin(ang);
    }
    return tw;
}

}  // namespace mr

namespace mr {
namespace {

// In-place radix-2 FFT of n = 2^k doubles: sign -1 forward, +1 inverse
// (unnormalised).
void fft_double(std::vector<std::complex<double>> &x, int sign)
{
    const size_t n = x.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    const double pi = 3.1415926535897932384626433832795;
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t h = len / 2;
        for (size_t k = 0; k < h; ++k) {
            const double ang = sign * 2.0 * pi * (double)k / (double)len;
            const std::complex<double> w(std::cos(ang), std::sin(ang));
            for (size_t i = k; i < n; i += len) {
                const std::complex<double> u = x[i], v = x[i + h] * w;
                x[i] = u + v;
                x[i + h] = u - v;
            }
        }
    }
}

}  // namespace

// The ProjectionFilterSingle filter of a row of width n as a short
// convolution.  The filter is linear in the padded line, which is the row at
// [ins, ins + n), its first value times the left taper and its last value
// times the right taper (ProjectionFilter.cpp:346-361).  Write the row as
// d + first (1 - s) + last s, s = i / (n - 1) in float (d is 0 at both
// ends, so its convolution has no edge jumps to cancel); then
//   y = conv(d, g) + first * resp_first + last * resp_last,
// with g = post0 * IFFT_P(hs) the filter's (real, even) impulse response,
// resp_first its response to the left taper plus 1 - s on the row and
// resp_last to the right taper plus s.  d reaches only lags |m| < n, so
// conv(d, g) is a Q-point circular convolution for any Q >= 2n - 1: the
// kernel runs a Q-point FFT (Q = P/2 for paddingdoubler 2) instead of the
// P-point one.  g, its Q-point spectrum (divided by Q) and the responses are
// computed here in double from the float s the kernel uses.  Returns Q, or 0
// when no power of two in [64, P) holds 2n - 1.
size_t short_filter(size_t P, size_t n, size_t ins, const std::vector<float> &hs, const std::vector<float> &taper,
                    float post0, std::vector<float> &hq, std::vector<float> &resp, std::vector<float> &resp_adj)
{
    size_t Q = 64;
    while (Q < 2 * n - 1) Q <<= 1;
    if (Q >= P) return 0;
    // g = post0 * IFFT_P(hs), unnormalised
    std::vector<std::complex<double>> g(P);
    for (size_t k = 0; k < P; ++k) g[k] = std::complex<double>((double)hs[k], 0.0);
    fft_double(g, +1);
    for (auto &v : g) v *= (double)post0;
    // the Q-point kernel: lags 0 .. n-1 and -(n-1) .. -1
    std::vector<std::complex<double>> gq(Q, std::complex<double>(0.0, 0.0));
    for (size_t m = 0; m < n; ++m) gq[m] = g[m];
    for (size_t m = 1; m < n; ++m) gq[Q - m] = g[P - m];
    fft_double(gq, -1);
    hq.resize(Q);
    for (size_t k = 0; k < Q; ++k) hq[k] = (float)(gq[k].real() / (double)Q);
    // responses at the output samples to the two tapers, exactly as the
    // padded line holds them (left: i < ins; right: P + 15 - i < ins,
    // i >= ins + n), with 1 - s (left) or s (right) on the row
    // (the adjoint's: the tapers alone)
    const float inv = 1.0f / (float)(n - 1);
    resp.assign(2 * n, 0.0f);
    resp_adj.assign(2 * n, 0.0f);
    for (int form = 0; form < 4; ++form) {
        const int side = form & 1;
        const bool ramp = form < 2;
        std::vector<std::complex<double>> t(P, std::complex<double>(0.0, 0.0));
        for (size_t i = 0; i < P; ++i) {
            if (side == 0 && i < ins) t[i] = (double)taper[i];
            if (side == 1 && i >= ins + n && P + 15 - i < ins) t[i] = (double)taper[P + 15 - i];
        }
        for (size_t i = 0; ramp && i < n; ++i) {
            const float s = (float)i * inv, t1 = 1.0f - s;
            t[ins + i] = (double)(side == 0 ? t1 : s);
        }
        fft_double(t, -1);
        for (size_t k = 0; k < P; ++k) t[k] *= (double)hs[k];
        fft_double(t, +1);
        std::vector<float> &dst = ramp ? resp : resp_adj;
        for (size_t i = 0; i < n; ++i) dst[side * n + i] = (float)(t[ins + i].real() * (double)post0);
    }
    return Q;
}

}  // namespace mr

// --------------------------------------------------------------------------
// ProjectionFilterSingle plan (C ABI)
// --------------------------------------------------------------------------
struct mr_filter_plan {
    int device = 0;
    size_t width = 0, P = 0, n_insert = 0;
    float scale = 0.0f;
    std::vector<float> H;       // P/2 (reference layout)
    float *d_hs = nullptr;      // P
    float *d_taper = nullptr;   // n_insert (>= 1)
    float2 *d_tw = nullptr;     // P
    // the short form (Q > 0): the spectrum of the Q-point convolution, its
    // twiddles, and the responses to the two tapers (2 x width)
    size_t Q = 0;
    float *d_hq = nullptr;
    float2 *d_twq = nullptr;
    float *d_resp = nullptr;    // 4 x width: forward (ramp) responses, then the adjoint's
};

extern "C" int mr_filter_plan_create(mr_filter_plan **plan, size_t width, int filter_type,
                                     float cutoff, float order, int use_bias, float bias_weight,
                                     size_t padding_doubler)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    *plan = nullptr;
    MR_REQUIRE(width >= 2, "filter width must be at least 2");
    MR_REQUIRE(filter_type >= 0 && filter_type <= 4, "unknown filter type");
    MR_REQUIRE(padding_doubler <= 8, "paddingdoubler out of range");
    std::vector<float> H, taper;
    size_t P = 0, n_insert = 0;
    int rc = mr::build_pb_filter(width, filter_type, cutoff, order, use_bias != 0, bias_weight,
                                 padding_doubler, P, n_insert, H, taper);
    if (rc != MR_OK) return rc;
    // the kernels run transforms of 64..16384 points: the P-point line, or
    // the short form's Q points (Q the power of two >= max(64, 2 width - 1)),
    // which also serves widths whose P exceeds 16384 (up to width 8192)
    size_t Qs = 64;
    while (Qs < 2 * width - 1) Qs <<= 1;
    const bool short_fits = Qs < P && Qs <= 16384;
    if (P < 64 || (P > 16384 && !short_fits) || P > ((size_t)1 << 20)) {
        mr::set_error("FFT length " + std::to_string(P) + " outside the supported 64..16384 (widths up to 8192 "
                      "through the short form)");
        return MR_ERR_UNSUPPORTED;
    }
    auto *p = new mr_filter_plan();
    p->width = width; p->P = P; p->n_insert = n_insert; p->H = H;
    p->scale = 3.1415926535897932384626f / (4.0f * (float)(P / 2));
    if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
    std::vector<float> hs(P);
    for (size_t k = 0; k < P; ++k)
        hs[k] = k < P / 2 ? H[k] : (k == P / 2 ? 1.0f : H[P - k]);
    std::vector<float> tw = mr::make_twiddles(P);
    if (taper.empty()) taper.push_back(0.0f);
    hipError_t e = hipMalloc(&p->d_hs, sizeof(float) * P);
    if (e == hipSuccess) e = hipMalloc(&p->d_taper, sizeof(float) * taper.size());
    if (e == hipSuccess) e = hipMalloc(&p->d_tw, sizeof(float2) * P);
    if (e == hipSuccess) e = hipMemcpy(p->d_hs, hs.data(), sizeof(float) * P, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_taper, taper.data(), sizeof(float) * taper.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_tw, tw.data(), sizeof(float) * 2 * P, hipMemcpyHostToDevice);
    // the short form of the same filter (a Q-point FFT), when Q < P
    std::vector<float> hq, resp, resp_adj;
    const size_t Q = (e == hipSuccess && short_fits)
                         ? mr::short_filter(P, width, n_insert, hs, taper, p->scale, hq, resp, resp_adj)
                         : 0;
    if (Q > 0) {
        const std::vector<float> twq = mr::make_twiddles(Q);
        resp.insert(resp.end(), resp_adj.begin(), resp_adj.end());
        e = hipMalloc(&p->d_hq, sizeof(float) * Q);
        if (e == hipSuccess) e = hipMalloc(&p->d_twq, sizeof(float2) * Q);
        if (e == hipSuccess) e = hipMalloc(&p->d_resp, sizeof(float) * 4 * width);
        if (e == hipSuccess) e = hipMemcpy(p->d_hq, hq.data(), sizeof(float) * Q, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->d_twq, twq.data(), sizeof(float) * 2 * Q, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(p->d_resp, resp.data(), sizeof(float) * 4 * width, hipMemcpyHostToDevice);
        p->Q = Q;
    }
    if (e != hipSuccess) {
        mr_filter_plan_destroy(p);
        return mr::hip_fail(e, "mr_filter_plan_create upload", __FILE__, __LINE__);
    }
    *plan = p;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_filter_plan_destroy(mr_filter_plan *plan)
{
    MR_GUARD_BEGIN
    if (!plan) return MR_OK;
    if (plan->d_hs) (void)hipFree(plan->d_hs);
    if (plan->d_taper) (void)hipFree(plan->d_taper);
    if (plan->d_tw) (void)hipFree(plan->d_tw);
    if (plan->d_hq) (void)hipFree(plan->d_hq);
    if (plan->d_twq) (void)hipFree(plan->d_twq);
    if (plan->d_resp) (void)hipFree(plan->d_resp);
    delete plan;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_filter_plan_info(const mr_filter_plan *plan, size_t *fft_size, size_t *n_insert,
                                   float *h_filter)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    if (fft_size) *fft_size = plan->P;
    if (n_insert) *n_insert = plan->n_insert;
    if (h_filter) std::copy(plan->H.begin(), plan->H.end(), h_filter);
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_filter_plan_kernel(const mr_filter_plan *plan, char *name, size_t len)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && name != nullptr && len > 0, "NULL argument");
    const size_t P = plan->Q > 0 ? plan->Q : plan->P;
    const size_t nt = std::min<size_t>(std::max<size_t>(P / 8, 64), 1024);
    const std::string d = "row_filter_kernel2<" + std::to_string(P) + "," + std::to_string(nt) + ">" +
                          (plan->Q > 0 ? " short" : "");
    std::snprintf(name, len, "%s", d.c_str());
    return MR_OK;
    MR_GUARD_END
}

int mr::filter_rows_to(mr_filter_plan *plan, const float *in, float *out, size_t n_rows, hipStream_t s)
{
    if (n_rows == 0) return MR_OK;
    mr::RowFilterArgs a{};
    a.rows = const_cast<float *>(in); a.n_rows = (long)n_rows; a.n = (int)plan->width; a.mode = 0;
    a.insert = (int)plan->n_insert; a.taper = plan->d_taper; a.hs = plan->d_hs; a.tw = plan->d_tw;
    a.post0 = plan->scale; a.post1 = 1.0f; a.post2 = 1.0f; a.pre_w = nullptr; a.rows_per_proj = 1;
    a.out = out;
    if (plan->Q > 0) {
        // the same filter as a Q-point convolution plus the taper responses
        a.mode = 3;
        a.hs = plan->d_hq;
        a.tw = plan->d_twq;
        a.resp_first = plan->d_resp;
        a.resp_last = plan->d_resp + plan->width;
        a.ramp_inv = 1.0f / (float)(plan->width - 1);
        return mr::run_row_filter(a, (int)plan->Q, s);
    }
    return mr::run_row_filter(a, (int)plan->P, s);
}

size_t mr::filter_width(const mr_filter_plan *plan) { return plan->width; }

extern "C" int mr_filter_rows_adjoint(mr_filter_plan *plan, float *d_rows, size_t n_rows, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(d_rows != nullptr || n_rows == 0, "rows pointer is NULL");
    if (n_rows == 0) return MR_OK;
    mr::RowFilterArgs a{};
    a.rows = d_rows; a.n_rows = (long)n_rows; a.n = (int)plan->width; a.mode = 2;
    a.insert = (int)plan->n_insert; a.taper = plan->d_taper; a.hs = plan->d_hs; a.tw = plan->d_tw;
    a.post0 = plan->scale; a.post1 = 1.0f; a.post2 = 1.0f; a.pre_w = nullptr; a.rows_per_proj = 1;
    a.out = nullptr;
    if (plan->Q > 0) {
        // adjoint of the short form: the same convolution (the impulse
        // response is even) plus the taper responses' dot products folded
        // onto the row ends
        a.mode = 4;
        a.hs = plan->d_hq;
        a.tw = plan->d_twq;
        a.resp_first = plan->d_resp + 2 * plan->width;
        a.resp_last = plan->d_resp + 3 * plan->width;
        return mr::run_row_filter(a, (int)plan->Q, (hipStream_t)stream);
    }
    return mr::run_row_filter(a, (int)plan->P, (hipStream_t)stream);
    MR_GUARD_END
}

extern "C" int mr_filter_rows_lognorm(mr_filter_plan *plan, float *d_proj, size_t n_proj, size_t rows,
                                      const float *d_flat_log, const float *d_dark, const float *dose,
                                      float dark_dose, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(n_proj == 0 || (d_proj != nullptr && rows > 0), "NULL projection pointer or empty projection");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (n_proj == 0) return MR_OK;
    hipStream_t s = (hipStream_t)stream;
    if (plan->Q == 0) {
        // the P-point filter has no fused load: normalise, then filter
        int rc = mr_lognorm(d_proj, n_proj, rows * plan->width, d_flat_log, d_dark, dose, dark_dose, stream);
        if (rc != MR_OK) return rc;
        return mr::filter_rows_to(plan, d_proj, d_proj, n_proj * rows, s);
    }
    // per-projection dose terms as mr_lognorm forms them
    float *d_dl = nullptr;
    std::vector<float> dl;
    if (dose) {
        dl.resize(n_proj);
        for (size_t j = 0; j < n_proj; ++j) {
            const float d = dose[j] - dark_dose;
            dl[j] = std::log(d < 1 ? 1.0f : d);
            MR_REQUIRE(dl[j] == dl[j], "The projection dose is a NaN");
        }
        MR_HIP(hipMallocAsync((void **)&d_dl, sizeof(float) * n_proj, s));
        hipError_t e = hipMemcpyAsync(d_dl, dl.data(), sizeof(float) * n_proj, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipFreeAsync(d_dl, s);
            return mr::hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
        }
    }
    mr::RowFilterArgs a{};
    a.rows = d_proj; a.n_rows = (long)(n_proj * rows); a.n = (int)plan->width; a.mode = 3;
    a.hs = plan->d_hq; a.tw = plan->d_twq; a.out = d_proj;
    a.resp_first = plan->d_resp; a.resp_last = plan->d_resp + plan->width;
    a.ramp_inv = 1.0f / (float)(plan->width - 1);
    a.rows_per_proj = (int)rows;
    a.norm = 1; a.norm_flat_log = d_flat_log; a.norm_dark = d_dark; a.norm_dl = d_dl;
    const int rc = mr::run_row_filter(a, (int)plan->Q, s);
    if (d_dl) (void)hipFreeAsync(d_dl, s);
    if (rc != MR_OK) return rc;
    // the host dose table must outlive the copy
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_filter_rows(mr_filter_plan *plan, float *d_rows, size_t n_rows, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr, "plan is NULL");
    MR_REQUIRE(d_rows != nullptr || n_rows == 0, "rows pointer is NULL");
    return mr::filter_rows_to(plan, d_rows, d_rows, n_rows, (hipStream_t)stream);
    MR_GUARD_END
}
