/*
 * CPU oracle for the MuhRec reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithms, used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker and as the timed CPU baseline.
 * No product code path links, loads or calls it.
 *
 * Two flavours of every function:
 *   *_f32 : "MuhRec-faithful" single precision.  Same float operation order,
 *           same truncations and the same quirks as the reference C++.  The
 *           only deviation is the FFT, which the reference takes from FFTW
 *           (float); here it is an exact double-precision DFT (radix-2),
 *           rounded to float where FFTW hands back floats.
 *   *_f64 : the same algorithm evaluated in double precision everywhere.
 *
 * Reference files restated (paths relative to the reference root):
 *   frameworks/tomography/Preprocessing/StdPreprocModules/src/ProjectionFilter.cpp
 *   core/kipl/kipl/src/fft/fftbasef.cpp
 *   core/kipl/kipl/include/math/core/mathfunctions.hpp
 *   frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp
 *   frameworks/tomography/Backprojectors/StdBackProjectors/src/StdBackProjectorBase.cpp
 *   frameworks/tomography/Backprojectors/StdBackProjectors/src/MultiProjBPparallel.cpp
 *   frameworks/tomography/Backprojectors/StdBackProjectors/src/MultiProjBP.cpp
 *   frameworks/tomography/Backprojectors/StdBackProjectors/src/NNMultiProjBP.cpp
 *   frameworks/tomography/Backprojectors/FDKBackProjectors/src/fdkreconbase.cpp
 *   frameworks/tomography/Backprojectors/FDKBackProjectors/src/fdkbp_single.cpp
 *   frameworks/tomography/Backprojectors/FDKBackProjectors/src/fdkbp.cpp
 *   frameworks/tomography/Preprocessing/StdPreprocModules/src/NormPlugins.cpp
 *   frameworks/tomography/Framework/ReconAlgorithms/ReconAlgorithms/linearforwardprojector.cpp
 *
 * Parity pinning of each function is documented beside its tests
 * (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FPI 3.1415926535897932384626f
#define ORC_DPI 3.1415926535897932384626433832795028841971693993751

/* ------------------------------------------------------------------------ */
/* double-precision complex FFT (radix-2, in place, power-of-two length)     */
/* ------------------------------------------------------------------------ */

typedef struct { double re, im; } orc_cplx;

static void orc_fft(orc_cplx *a, size_t n, int sign, const orc_cplx *tw /* n/2 */)
{
    /* bit reversal */
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { orc_cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < half; ++k) {
                orc_cplx w = tw[k * step];
                if (sign > 0) w.im = -w.im;
                orc_cplx u = a[i + k], v = a[i + k + half];
                orc_cplx t = { v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re };
                a[i + k].re = u.re + t.re; a[i + k].im = u.im + t.im;
                a[i + k + half].re = u.re - t.re; a[i + k + half].im = u.im - t.im;
            }
        }
    }
}

/* twiddles exp(-2 pi i k / n), k < n/2 */
static orc_cplx *orc_twiddles(size_t n)
{
    orc_cplx *tw = (orc_cplx *)malloc(sizeof(orc_cplx) * (n / 2 + 1));
    for (size_t k = 0; k < n / 2 + 1; ++k) {
        double ang = -2.0 * ORC_DPI * (double)k / (double)n;
        tw[k].re = cos(ang); tw[k].im = sin(ang);
    }
    return tw;
}

/* exposed for the oracle's own self-test against a brute-force DFT */
int orc_fft_c2c(double *re, double *im, size_t n, int sign)
{
    if (n == 0 || (n & (n - 1))) return -1;
    orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * n);
    orc_cplx *tw = orc_twiddles(n);
    for (size_t i = 0; i < n; ++i) { a[i].re = re[i]; a[i].im = im[i]; }
    orc_fft(a, n, sign, tw);
    for (size_t i = 0; i < n; ++i) { re[i] = a[i].re; im[i] = a[i].im; }
    free(a); free(tw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ProjectionFilterSingle (SPP/src/ProjectionFilter.cpp)                     */
/* ------------------------------------------------------------------------ */

/* ComputeFilterSize, ProjectionFilter.cpp:120-125 */
size_t orc_filter_size(size_t len, size_t paddingdoubler)
{
    double e = log((double)len) / log(2.0);
    return (size_t)1 << ((size_t)ceil(e) + paddingdoubler);
}

enum { ORC_RAMLAK = 0, ORC_SHEPPLOGAN, ORC_HANNING, ORC_HAMMING, ORC_BUTTERWORTH };

/* ProjectionFilterSingle::BuildFilter, ProjectionFilter.cpp:254-302.
 * Returns 0, or -1 when the cut-off is outside (0, 0.5] (the reference throws). */
int orc_filter_build_f32(int ftype, float cutoff, float order, int usebias, float biasweight,
                         size_t P, float *H)
{
    const size_t N2 = P / 2;
    if (!((0 < cutoff) && (cutoff <= 0.5f))) return -1;
    float FilterOrder = 0.0f;
    if (ftype == ORC_BUTTERWORTH) FilterOrder = order;
    const size_t c = (size_t)(2 * cutoff * (float)N2);
    float rampstep = 1.0f / (((float)N2 - 1.0f));
    for (size_t i = 0; i < N2; ++i) H[i] = (float)i * rampstep;
    float fstep = 1.0f / (((float)c - 1.0f));
    for (size_t i = 0; i < c && i < N2; ++i) {
        switch (ftype) {
        case ORC_RAMLAK: break;
        case ORC_SHEPPLOGAN:
            H[i] = (float)(sin(ORC_DPI * (double)i * (double)fstep * (double)cutoff)) / (float)N2;
            break;
        case ORC_HANNING:
            H[i] *= (float)(0.5f + 0.5f * cosf((float)i * fstep * ORC_FPI * cutoff));
            break;
        case ORC_HAMMING:
            H[i] *= (float)(0.54 + 0.46 * cos((double)((float)i * fstep) * ORC_DPI * (double)cutoff));
            break;
        case ORC_BUTTERWORTH:
            H[i] /= (float)(1 + powf((float)i * fstep * cutoff, FilterOrder));
            break;
        default: break;
        }
    }
    for (size_t i = c; i < N2; ++i) H[i] = 0.0f;
    if (usebias) H[0] = biasweight * H[1];
    return 0;
}

/* Same filter, every quantity in double. */
int orc_filter_build_f64(int ftype, double cutoff, double order, int usebias, double biasweight,
                         size_t P, double *H)
{
    const size_t N2 = P / 2;
    if (!((0 < cutoff) && (cutoff <= 0.5))) return -1;
    double FilterOrder = (ftype == ORC_BUTTERWORTH) ? order : 0.0;
    const size_t c = (size_t)(2 * cutoff * (double)N2);
    double rampstep = 1.0 / ((double)N2 - 1.0);
    for (size_t i = 0; i < N2; ++i) H[i] = (double)i * rampstep;
    double fstep = 1.0 / ((double)c - 1.0);
    for (size_t i = 0; i < c && i < N2; ++i) {
        double t = (double)i * fstep * cutoff;
        switch (ftype) {
        case ORC_SHEPPLOGAN: H[i] = sin(ORC_DPI * t) / (double)N2; break;
        case ORC_HANNING: H[i] *= 0.5 + 0.5 * cos(ORC_DPI * t); break;
        case ORC_HAMMING: H[i] *= 0.54 + 0.46 * cos(ORC_DPI * t); break;
        case ORC_BUTTERWORTH: H[i] /= 1 + pow(t, FilterOrder); break;
        default: break;
        }
    }
    for (size_t i = c; i < N2; ++i) H[i] = 0.0;
    if (usebias) H[0] = biasweight * H[1];
    return 0;
}

/* PreparePadding, ProjectionFilter.cpp:363-375 with Sigmoid from
 * core/kipl/kipl/include/math/core/mathfunctions.hpp:60-63 (float instance:
 * exp takes a float argument).  Returns nInsert. */
size_t orc_filter_taper_f32(size_t nImage, size_t P, float *taper)
{
    size_t nInsert = (P / 2) - (nImage / 2);
    for (size_t i = 0; i < nInsert; ++i) {
        float x = (float)i / (float)nInsert - 0.5f;
        taper[i] = (float)(1.0 / (1.0 + expf(-(x - 0.0f) / 0.07f)));
    }
    return nInsert;
}

size_t orc_filter_taper_f64(size_t nImage, size_t P, double *taper)
{
    size_t nInsert = (P / 2) - (nImage / 2);
    for (size_t i = 0; i < nInsert; ++i) {
        double x = (double)i / (double)nInsert - 0.5;
        taper[i] = 1.0 / (1.0 + exp(-x / 0.07));
    }
    return nInsert;
}

/* Pad, ProjectionFilter.cpp:346-361: destination of length P+16; only the
 * first P samples enter the FFT. */
static void orc_pad_f64(const float *src, size_t n, double *dst, size_t P,
                        const double *tap, size_t nInsert, int f32)
{
    const size_t nDest = P + 16;
    for (size_t i = 0; i < P; ++i) dst[i] = 0.0;
    for (size_t i = 0; i < n; ++i) dst[nInsert + i] = src[i];
    for (size_t i = 0; i < nInsert; ++i) {
        double l, r;
        if (f32) { l = (double)(src[0] * (float)tap[i]); r = (double)(src[n - 1] * (float)tap[i]); }
        else     { l = (double)src[0] * tap[i];           r = (double)src[n - 1] * tap[i]; }
        dst[i] = l;
        size_t j = nDest - 1 - i;
        if (j < P) dst[j] = r;
    }
}

/* FilterProjection, ProjectionFilter.cpp:304-344.  rows: nrows x n, row-major,
 * filtered in place.  Bins k < P/2 are scaled by H[k], the Nyquist bin P/2 is
 * passed through, output sample i = c2r(...)[i+nInsert] * pi/(2P). */
void orc_filter_rows_f32(float *rows, size_t nrows, size_t n, size_t P,
                         const float *H, const float *taper, size_t nInsert)
{
    const size_t N2 = P / 2;
    orc_cplx *tw = orc_twiddles(P);
    double *tapd = (double *)malloc(sizeof(double) * (nInsert + 1));
    for (size_t i = 0; i < nInsert; ++i) tapd[i] = taper[i];
    const float scale = ORC_FPI / (4.0f * (float)N2);
#pragma omp parallel
    {
        orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * P);
        double *pad = (double *)malloc(sizeof(double) * P);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nrows; ++r) {
            float *row = rows + (size_t)r * n;
            orc_pad_f64(row, n, pad, P, tapd, nInsert, 1);
            for (size_t i = 0; i < P; ++i) { a[i].re = pad[i]; a[i].im = 0.0; }
            orc_fft(a, P, -1, tw);
            /* FFTW hands back float spectra; the filter multiplies in float. */
            for (size_t k = 0; k <= N2; ++k) {
                float re = (float)a[k].re, im = (float)a[k].im;
                if (k < N2) { re *= H[k]; im *= H[k]; }
                a[k].re = re; a[k].im = (k == 0 || k == N2) ? 0.0 : im;
            }
            for (size_t k = 1; k < N2; ++k) { a[P - k].re = a[k].re; a[P - k].im = -a[k].im; }
            orc_fft(a, P, +1, tw);
            for (size_t i = 0; i < n; ++i) row[i] = (float)a[i + nInsert].re * scale;
        }
        free(a); free(pad);
    }
    free(tapd); free(tw);
}

/* Same filter in double precision: in (float rows) -> out (double rows). */
void orc_filter_rows_f64(const float *in, double *out, size_t nrows, size_t n, size_t P,
                         const double *H, const double *taper, size_t nInsert)
{
    const size_t N2 = P / 2;
    orc_cplx *tw = orc_twiddles(P);
    const double scale = ORC_DPI / (2.0 * (double)P);
#pragma omp parallel
    {
        orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * P);
        double *pad = (double *)malloc(sizeof(double) * P);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nrows; ++r) {
            orc_pad_f64(in + (size_t)r * n, n, pad, P, taper, nInsert, 0);
            for (size_t i = 0; i < P; ++i) { a[i].re = pad[i]; a[i].im = 0.0; }
            orc_fft(a, P, -1, tw);
            for (size_t k = 0; k <= N2; ++k) {
                double h = (k < N2) ? H[k] : 1.0;
                a[k].re *= h; a[k].im = (k == 0 || k == N2) ? 0.0 : a[k].im * h;
            }
            for (size_t k = 1; k < N2; ++k) { a[P - k].re = a[k].re; a[P - k].im = -a[k].im; }
            orc_fft(a, P, +1, tw);
            for (size_t i = 0; i < n; ++i) out[(size_t)r * n + i] = a[i + nInsert].re * scale;
        }
        free(a); free(pad);
    }
    free(tw);
}

/* Adjoint of the filter (the project's own operator for gradients; the
 * reference has none).  The forward is out = scale * M pad(x) restricted to
 * [nInsert, nInsert+n), with M = c2r(Hs * r2c(.)) symmetric (Hs real and
 * even), so the adjoint embeds g at nInsert, applies scale * M, and folds
 * the two tapers back onto the first and last sample. */
void orc_filter_rows_adjoint_f64(const double *g, double *out, size_t nrows, size_t n, size_t P,
                                 const double *H, const double *taper, size_t nInsert)
{
    const size_t N2 = P / 2, nDest = P + 16;
    orc_cplx *tw = orc_twiddles(P);
    const double scale = ORC_DPI / (2.0 * (double)P);
#pragma omp parallel
    {
        orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * P);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nrows; ++r) {
            for (size_t i = 0; i < P; ++i) { a[i].re = 0.0; a[i].im = 0.0; }
            for (size_t i = 0; i < n; ++i) a[nInsert + i].re = g[(size_t)r * n + i] * scale;
            orc_fft(a, P, -1, tw);
            for (size_t k = 0; k <= N2; ++k) {
                double h = (k < N2) ? H[k] : 1.0;
                a[k].re *= h; a[k].im = (k == 0 || k == N2) ? 0.0 : a[k].im * h;
            }
            for (size_t k = 1; k < N2; ++k) { a[P - k].re = a[k].re; a[P - k].im = -a[k].im; }
            orc_fft(a, P, +1, tw);
            double *o = out + (size_t)r * n;
            for (size_t i = 0; i < n; ++i) o[i] = a[nInsert + i].re;
            for (size_t i = 0; i < nInsert; ++i) {
                o[0] += taper[i] * a[i].re;
                const size_t j = nDest - 1 - i;
                if (j < P) o[n - 1] += taper[i] * a[j].re;
            }
        }
        free(a);
    }
    free(tw);
}

/* ------------------------------------------------------------------------ */
/* Circle mask, RF/src/BackProjectorModuleBase.cpp:153-218.                 */
/* roi == NULL: no MatrixInfo ROI.  roi = {x0,y0,x1,y1} inclusive,           */
/* cone != 0 selects the cone-beam ROI clipping branch.  shrink: the radius  */
/* reduction of the tilt correction (orc_tilt_mask_shrink; 0 without tilt).  */
/* ------------------------------------------------------------------------ */
void orc_circle_mask_ex(size_t nSizeX, size_t nSizeY, const size_t *roi, int cone, float shrink,
                        size_t *first, size_t *second)
{
    const float cx = (float)(nSizeX >> 1);
    const float cy = (float)(nSizeY >> 1);
    float R = cx - 1;
    R -= shrink;
    const float R2 = R * R;
    for (size_t i = 0; i < nSizeY; ++i) {
        float y = (float)i - cy;
        float y2 = y * y;
        if (y2 <= R2) {
            float x = sqrtf(R2 - y2);
            first[i] = (size_t)ceilf(cx - x);
            size_t s = (size_t)floorf(cx + x);
            second[i] = (nSizeX - 1u) < s ? (nSizeX - 1u) : s;
        } else {
            first[i] = 0u; second[i] = 0u;
        }
        if (roi) {
            if (roi[1] <= i && i <= roi[3]) {
                if (!cone) {
                    if (first[i] < roi[0]) first[i] = roi[0];
                    if (second[i] > roi[2]) second[i] = roi[2];
                } else {
                    size_t a = nSizeX - roi[2], b = nSizeX - roi[0];
                    if (first[i] < a) first[i] = a;
                    if (second[i] > b) second[i] = b;
                }
            } else {
                first[i] = 0u; second[i] = 0u;
            }
        }
    }
}

void orc_circle_mask(size_t nSizeX, size_t nSizeY, const size_t *roi, int cone,
                     size_t *first, size_t *second)
{
    orc_circle_mask_ex(nSizeX, nSizeY, roi, cone, 0.0f, first, second);
}

/* BackProjectorModuleBase.cpp:171-180: with bCorrectTilt the mask radius
 * shrinks by floor(tan(|fTiltAngle|) * slices), slices = ProjectionInfo.nDims[1]
 * (roi[3] for repeated sinograms); float operations (the kipl headers bring
 * "using namespace std", so fabs/tan/floor take their float overloads). */
float orc_tilt_mask_shrink(float tilt_angle, size_t slices)
{
    return floorf(tanf(fabsf(tilt_angle) * ORC_FPI / 180.0f) * (float)slices);
}

/* ------------------------------------------------------------------------ */
/* Parallel-beam back-projection                                             */
/* StdBackProjectorBase::Process(2D/3D), SBP/src/StdBackProjectorBase.cpp:64-149 */
/* MultiProjectionBPparallel::BackProject, SBP/src/MultiProjBPparallel.cpp:115-167 */
/* MultiProjectionBP::BackProject, SBP/src/MultiProjBP.cpp:95-143            */
/* NearestNeighborBP::BackProject, SBP/src/NNMultiProjBP.cpp:80-118          */
/* ------------------------------------------------------------------------ */

/* Indices of the projections that the reference back-projects, in order.
 * bufsize = ProjectionBufferSize; 0 disables the buffer drop (every
 * projection used).  Returns the count. */
size_t orc_used_projections(size_t np, size_t bufsize, size_t *used)
{
    size_t cnt = 0, c = 0;
    size_t slot[4096];
    for (size_t i = 0; i < np; ++i) {
        if (bufsize == 0) { used[cnt++] = i; continue; }
        slot[c] = i; c++;
        int last = (i == np - 1);
        if (last || bufsize <= c) {
            if (bufsize <= c) c--;
            for (size_t k = 0; k < c; ++k) used[cnt++] = slot[k];
            c = 0;
        }
    }
    return cnt;
}

/* Per-projection tables, StdBackProjectorBase.cpp:73-80 (float sin/cos of a
 * float argument). */
void orc_bp_tables_f32(const float *angles, size_t np, int direction, float center,
                       size_t matrixCenterX, float *fsin, float *fcos, float *fstartu)
{
    float dirWeight = 2.0f * ((float)direction - 0.5f);
    for (size_t i = 0; i < np; ++i) {
        fsin[i] = sinf(dirWeight * angles[i] * ORC_FPI / 180.0f);
        fcos[i] = cosf(dirWeight * angles[i] * ORC_FPI / 180.0f);
        fstartu[i] = (float)matrixCenterX * (fsin[i] - fcos[i]) + center;
    }
}

/* proj: np x nv x nu (u fastest, one projection after the other, already
 *       filtered, NOT yet weighted).
 * vol : nu(y) x nu(x) x nv(z), z fastest (MatrixZXY), overwritten.
 * interp: 0 linear (MultiProjBPparallel / MultiProjBP), 1 nearest (NearestNeighborBP)
 * multiprojbp: 1 reproduces MultiProjBP's B-column offset 4*(nv/4) instead of
 *       the padded column length (identical when nv % 4 == 0).
 * Only slices z < 4*(nv/4) are updated (SizeZ = volume.Size(0)/4). */
void orc_bp_parallel_f32(const float *proj, size_t nu, size_t nv, size_t np,
                         const float *angles, const float *weights, float center, int direction,
                         int interp, size_t bufsize, int multiprojbp,
                         const size_t *mfirst, const size_t *msecond, float *vol)
{
    const size_t vpad = (nv + 3) & ~(size_t)3;
    const size_t sizeZ4 = (nv / 4) * 4;
    const size_t bstride = multiprojbp ? (nv / 4) * 4 : vpad;
    const int sizeUm2 = (int)nu - 2;
    size_t *used = (size_t *)malloc(sizeof(size_t) * (np + 1));
    size_t nused = orc_used_projections(np, bufsize, used);
    float *fsin = (float *)malloc(sizeof(float) * np);
    float *fcos = (float *)malloc(sizeof(float) * np);
    float *fsu = (float *)malloc(sizeof(float) * np);
    orc_bp_tables_f32(angles, np, direction, center, nu / 2, fsin, fcos, fsu);
    /* weighted, transposed projection buffer: [nused][nu][vpad] (+1 column guard) */
    float *buf = (float *)calloc((nused * nu + 1) * vpad + 8, sizeof(float));
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)nused; ++k) {
        const float *p = proj + used[k] * nu * nv;
        float w = weights[used[k]];
        for (size_t v = 0; v < nv; ++v)
            for (size_t u = 0; u < nu; ++u)
                buf[((size_t)k * nu + u) * vpad + v] = p[v * nu + u] * w;
    }
    memset(vol, 0, sizeof(float) * nu * nu * nv);
    /* projections in chunks of 16, each back-projected over the whole slice
     * block before the next (the reference's ProjectionBufferSize loop,
     * StdBackProjectorBase.cpp:104-114, MultiProjBPparallel.cpp:115-167):
     * the chunk's columns stay in cache; every voxel still adds its
     * projections in increasing order */
    const size_t CH = 16;
    for (size_t k0 = 0; k0 < nused; k0 += CH) {
    const size_t k1 = nused < k0 + CH ? nused : k0 + CH;
#pragma omp parallel
    {
        float lsu[16];
#pragma omp for schedule(dynamic, 4)
        for (long y = 1; y <= (long)nu; ++y) {
            const size_t x0 = mfirst[y - 1], x1 = msecond[y - 1];
            for (size_t k = k0; k < k1; ++k) lsu[k - k0] = fsu[used[k]] + fcos[used[k]] * (float)y;
            for (size_t x = x0 + 1; x <= x1; ++x) {
                float *col = vol + ((size_t)(y - 1) * nu + (x - 1)) * nv;
                for (size_t k = k0; k < k1; ++k) {
                    float u = lsu[k - k0] - fsin[used[k]] * (float)x;
                    int n = (int)u;
                    if ((n < 0) || (sizeUm2 < n)) continue;
                    const float *A = buf + ((size_t)k * nu + (size_t)n) * vpad;
                    if (interp == 1) {
                        for (size_t z = 0; z < sizeZ4; ++z) col[z] = col[z] + A[z];
                    } else {
                        const float *B = A + bstride;
                        const float wb = fabsf(u - (float)n);
                        const float wa = 1.0f - wb;
                        for (size_t z = 0; z < sizeZ4; ++z) {
                            float a = A[z] * wa;
                            float b = B[z] * wb;
                            float s = a + b;
                            col[z] = col[z] + s;
                        }
                    }
                }
            }
        }
    }
    }
    free(buf); free(used); free(fsin); free(fcos); free(fsu);
}

/* Axis-tilt terms of the three parallel-beam back-projectors
 * (MultiProjBPparallel.cpp:68-75, MultiProjBP.cpp:63-64, NNMultiProjBP.cpp:61-62):
 * centeroffset = (roi[1] - fTiltPivotPosition) * tan(fTiltAngle pi / 180) and
 * centerinc = 4 tan(fTiltAngle pi / 180), in float (tan of a float argument
 * resolves to the float overload); row0 = ProjectionInfo.roi[1] of the block. */
void orc_bp_tilt_terms(float tilt_angle, float tilt_pivot, size_t row0, float *centeroffset, float *centerinc)
{
    const float t = tanf(tilt_angle * ORC_FPI / 180.0f);
    *centeroffset = ((float)row0 - tilt_pivot) * t;
    *centerinc = 4.0f * t;
}

/* The bCorrectTilt branches of the back-projectors, as written in the
 * reference (their weights are not a physical tilt model; they are
 * reproduced as they are):
 *   variant 0, MultiProjBPparallel (MultiProjBPparallel.cpp:49-114):
 *     u = ((fStartU + fCos y) - centeroffset) - fSin x, n = (int)u, f = u - n;
 *     slice group z (4 slices) weights B by |f_z - z centerinc|, where f_z is
 *     f with centerinc subtracted z times (fPosU -= centerinc after each group);
 *   variant 1, MultiProjBP (MultiProjBP.cpp:50-96):
 *     u = ((fStartU + fCos y) - fSin x) - centeroffset, n = (int)u;
 *     group z weights B by |(u - n) - z centerinc|;
 *   variant 2, NearestNeighborBP (NNMultiProjBP.cpp:48-81): u as variant 1,
 *     column n added unweighted.
 * In every variant the two columns read are n and n + 1 (B at the padded
 * column length for variant 0, 4*(nv/4) for variant 1) for all groups, and
 * interpA = 1 - interpB.  Arguments as orc_bp_parallel_f32 (proj not yet
 * weighted; vol ZXY, overwritten); the mask is the caller's (with the tilt
 * shrink, orc_tilt_mask_shrink). */
void orc_bp_parallel_tilt_f32(const float *proj, size_t nu, size_t nv, size_t np,
                              const float *angles, const float *weights, float center, int direction,
                              int variant, size_t bufsize, float tilt_angle, float tilt_pivot, size_t row0,
                              const size_t *mfirst, const size_t *msecond, float *vol)
{
    const size_t vpad = (nv + 3) & ~(size_t)3;
    const size_t sizeZ = nv / 4;
    const size_t bstride = variant == 1 ? sizeZ * 4 : vpad;
    const int sizeUm2 = (int)nu - 2;
    float co, ci;
    orc_bp_tilt_terms(tilt_angle, tilt_pivot, row0, &co, &ci);
    size_t *used = (size_t *)malloc(sizeof(size_t) * (np + 1));
    size_t nused = orc_used_projections(np, bufsize, used);
    float *fsin = (float *)malloc(sizeof(float) * np);
    float *fcos = (float *)malloc(sizeof(float) * np);
    float *fsu = (float *)malloc(sizeof(float) * np);
    orc_bp_tables_f32(angles, np, direction, center, nu / 2, fsin, fcos, fsu);
    float *buf = (float *)calloc((nused * nu + 1) * vpad + 8, sizeof(float));
    for (size_t k = 0; k < nused; ++k) {
        const float *p = proj + used[k] * nu * nv;
        float w = weights[used[k]];
        for (size_t v = 0; v < nv; ++v)
            for (size_t u = 0; u < nu; ++u)
                buf[(k * nu + u) * vpad + v] = p[v * nu + u] * w;
    }
    memset(vol, 0, sizeof(float) * nu * nu * nv);
#pragma omp parallel for schedule(dynamic, 4)
    for (long y = 1; y <= (long)nu; ++y) {
        const size_t x0 = mfirst[y - 1], x1 = msecond[y - 1];
        for (size_t x = x0 + 1; x <= x1; ++x) {
            float *col = vol + ((size_t)(y - 1) * nu + (x - 1)) * nv;
            for (size_t k = 0; k < nused; ++k) {
                const size_t i = used[k];
                float u;
                if (variant == 0) {
                    const float lsu = fsu[i] + fcos[i] * (float)y - co;
                    u = lsu - fsin[i] * (float)x;
                } else {
                    const float lsu = fsu[i] + fcos[i] * (float)y;
                    u = lsu - fsin[i] * (float)x - co;
                }
                const int n = (int)u;
                if ((n < 0) || (sizeUm2 < n)) continue;
                const float *A = buf + (k * nu + (size_t)n) * vpad;
                const float *B = A + bstride;
                if (variant == 2) {
                    for (size_t z = 0; z < 4 * sizeZ; ++z) col[z] = col[z] + A[z];
                    continue;
                }
                float f = u - (float)n;
                for (size_t z = 0; z < sizeZ; ++z) {
                    const float wb = variant == 0 ? fabsf(f - (float)z * ci) : fabsf(u - (float)n - (float)z * ci);
                    const float wa = 1.0f - wb;
                    for (size_t c = 4 * z; c < 4 * z + 4; ++c) {
                        float a = A[c] * wa;
                        float b = B[c] * wb;
                        float sm = a + b;
                        col[c] = col[c] + sm;
                    }
                    f -= ci;
                }
            }
        }
    }
    free(buf); free(used); free(fsin); free(fcos); free(fsu);
}

/* Adjoint of orc_bp_parallel_f32 with multiprojbp = 0 (MultiProjBPparallel,
 * NearestNeighborBP): the float positions n, w of the forward, sums in
 * double.  g: volume-shaped [nu(y)][nu(x)][nv]; out: [np][nv][nu], zero for
 * dropped projections and for slices z >= 4*(nv/4). */
void orc_bp_parallel_adjoint_f64(const double *g, size_t nu, size_t nv, size_t np,
                                 const float *angles, const float *weights, float center, int direction,
                                 int interp, size_t bufsize, const size_t *mfirst, const size_t *msecond,
                                 double *out)
{
    const size_t sizeZ4 = (nv / 4) * 4;
    const int sizeUm2 = (int)nu - 2;
    size_t *used = (size_t *)malloc(sizeof(size_t) * (np + 1));
    size_t nused = orc_used_projections(np, bufsize, used);
    float *fsin = (float *)malloc(sizeof(float) * np);
    float *fcos = (float *)malloc(sizeof(float) * np);
    float *fsu = (float *)malloc(sizeof(float) * np);
    orc_bp_tables_f32(angles, np, direction, center, nu / 2, fsin, fcos, fsu);
    memset(out, 0, sizeof(double) * np * nv * nu);
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < (long)nused; ++k) {
        const size_t p = used[k];
        const double wk = (double)weights[p];
        double *o = out + p * nv * nu;
        for (long y = 1; y <= (long)nu; ++y) {
            const float lsu = fsu[p] + fcos[p] * (float)y;
            for (size_t x = mfirst[y - 1] + 1; x <= msecond[y - 1]; ++x) {
                const float u = lsu - fsin[p] * (float)x;
                const int nn = (int)u;
                if ((nn < 0) || (sizeUm2 < nn)) continue;
                const double *col = g + ((size_t)(y - 1) * nu + (x - 1)) * nv;
                if (interp == 1) {
                    for (size_t z = 0; z < sizeZ4; ++z) o[z * nu + nn] += wk * col[z];
                } else {
                    const float wb = fabsf(u - (float)nn);
                    const float wa = 1.0f - wb;
                    for (size_t z = 0; z < sizeZ4; ++z) {
                        o[z * nu + nn] += wk * (double)wa * col[z];
                        o[z * nu + nn + 1] += wk * (double)wb * col[z];
                    }
                }
            }
        }
    }
    free(used); free(fsin); free(fcos); free(fsu);
}

/* The same back-projection in double precision (tables, positions,
 * interpolation and accumulation). */
void orc_bp_parallel_f64(const double *proj, size_t nu, size_t nv, size_t np,
                         const double *angles, const double *weights, double center, int direction,
                         int interp, size_t bufsize,
                         const size_t *mfirst, const size_t *msecond, double *vol)
{
    const size_t sizeZ4 = (nv / 4) * 4;
    const int sizeUm2 = (int)nu - 2;
    size_t *used = (size_t *)malloc(sizeof(size_t) * (np + 1));
    size_t nused = orc_used_projections(np, bufsize, used);
    double *fsin = (double *)malloc(sizeof(double) * np);
    double *fcos = (double *)malloc(sizeof(double) * np);
    double *fsu = (double *)malloc(sizeof(double) * np);
    double dirWeight = 2.0 * ((double)direction - 0.5);
    double mcx = (double)(nu / 2);
    for (size_t i = 0; i < np; ++i) {
        fsin[i] = sin(dirWeight * angles[i] * ORC_DPI / 180.0);
        fcos[i] = cos(dirWeight * angles[i] * ORC_DPI / 180.0);
        fsu[i] = mcx * (fsin[i] - fcos[i]) + center;
    }
    memset(vol, 0, sizeof(double) * nu * nu * nv);
#pragma omp parallel for schedule(dynamic, 4)
    for (long y = 1; y <= (long)nu; ++y) {
        const size_t x0 = mfirst[y - 1], x1 = msecond[y - 1];
        for (size_t x = x0 + 1; x <= x1; ++x) {
            double *col = vol + ((size_t)(y - 1) * nu + (x - 1)) * nv;
            for (size_t k = 0; k < nused; ++k) {
                size_t i = used[k];
                double u = fsu[i] + fcos[i] * (double)y - fsin[i] * (double)x;
                int n = (int)u;
                if ((n < 0) || (sizeUm2 < n)) continue;
                const double *p = proj + i * nu * nv;
                double w = weights[i];
                if (interp == 1) {
                    for (size_t z = 0; z < sizeZ4; ++z) col[z] += w * p[z * nu + n];
                } else {
                    double wb = fabs(u - (double)n), wa = 1.0 - wb;
                    for (size_t z = 0; z < sizeZ4; ++z)
                        col[z] += w * (p[z * nu + n] * wa + p[z * nu + n + 1] * wb);
                }
            }
        }
    }
    free(used); free(fsin); free(fcos); free(fsu);
}

/* ------------------------------------------------------------------------ */
/* FDK cone-beam (FDKbp_single), FDK/fdkreconbase.cpp:217-277 and           */
/* FDK/fdkbp_single.cpp:83-531, 641-683, 832-982                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    float sod, sdd;          /* ProjectionInfo.fSOD, fSDD */
    float pix[2];            /* ProjectionInfo.fResolution */
    float vox[3];            /* MatrixInfo.fVoxelSize */
    float ppoint[2];         /* ProjectionInfo.fpPoint */
    float center;            /* ProjectionInfo.fCenter */
    int direction;           /* 0 CW, 1 CCW */
    uint64_t roi[4];         /* ProjectionInfo.roi at back-projection time (slab) */
    uint64_t proj_roi3;      /* ProjectionInfo.projection_roi[3] */
    int correct_tilt;        /* ProjectionInfo.bCorrectTilt */
    float tilt_angle;        /* ProjectionInfo.fTiltAngle (degrees) */
    float tilt_pivot;        /* ProjectionInfo.fTiltPivotPosition */
} orc_fdk_geom;

/* getProjMatrix, fdkbp_single.cpp:83-255 (tgt = 0, vup = z). */
void orc_fdk_projmatrix_f32(const orc_fdk_geom *g, float angle, float *pm /*12*/)
{
    float ext[16] = {0}, intr[12] = {0};
    float sad = g->sod, sid = g->sdd;
    float offCenter = ((g->ppoint[0] - (float)g->roi[0]) - g->center) * g->vox[0];
    float cam[3] = {0.0f, 0.0f, 0.0f};
    if (g->direction == 0) {
        cam[0] = sad * cosf(-angle * ORC_FPI / 180.0f);
        cam[1] = sad * sinf(-angle * ORC_FPI / 180.0f);
    } else {
        cam[0] = sad * cosf(angle * ORC_FPI / 180.0f);
        cam[1] = sad * sinf(angle * ORC_FPI / 180.0f);
    }
    float vup[3] = {0.0f, 0.0f, 1.0f}, tgt[3] = {0.0f, 0.0f, 0.0f};
    float nrm[3], plt[3], pup[3];
    nrm[0] = cam[0] - tgt[0]; nrm[1] = cam[1] - tgt[1]; nrm[2] = cam[2] - tgt[2];
    float nn = sqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    nrm[0] /= nn; nrm[1] /= nn; nrm[2] /= nn;
    plt[0] = nrm[1] * vup[2] - nrm[2] * vup[1];
    plt[1] = nrm[2] * vup[0] - nrm[0] * vup[2];
    plt[2] = nrm[0] * vup[1] - nrm[1] * vup[0];
    float np_ = sqrtf(plt[0] * plt[0] + plt[1] * plt[1] + plt[2] * plt[2]);
    plt[0] /= np_; plt[1] /= np_; plt[2] /= np_;
    pup[0] = plt[1] * nrm[2] - plt[2] * nrm[1];
    pup[1] = plt[2] * nrm[0] - plt[0] * nrm[2];
    pup[2] = plt[0] * nrm[1] - plt[1] * nrm[0];
    float nu_ = sqrtf(pup[0] * pup[0] + pup[1] * pup[1] + pup[2] * pup[2]);
    pup[0] /= nu_; pup[1] /= nu_; pup[2] /= nu_;
    ext[0] = -1.0f * plt[0]; ext[1] = -1.0f * plt[1]; ext[2] = -1.0f * plt[2];
    ext[4] = -1.0f * pup[0]; ext[5] = -1.0f * pup[1]; ext[6] = -1.0f * pup[2];
    ext[8] = -1.0f * nrm[0]; ext[9] = -1.0f * nrm[1]; ext[10] = -1.0f * nrm[2];
    ext[15] = 1.0f;
    ext[3] = plt[0] * tgt[0] + plt[1] * tgt[1] + plt[2] * tgt[2] - offCenter;
    ext[7] = pup[0] * tgt[0] + pup[1] * tgt[1] + pup[2] * tgt[2];
    ext[11] = nrm[0] * tgt[0] + nrm[1] * tgt[1] + nrm[2] * tgt[2] + sad;
    intr[0] = 1.0f / g->pix[0];
    intr[5] = 1.0f / g->pix[1];
    intr[10] = 1.0f / sid;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += intr[i * 4 + k] * ext[k * 4 + j];
            pm[i * 4 + j] = acc;
        }
}

/* Detector-row window, fdkbp_single.cpp:322-367 (identical to the engine's
 * CBCT_roi, RF/src/ReconEngine.cpp:954-1008).  nvol = volume.Size(1). */
static void orc_fdk_cbct_roi_impl(const orc_fdk_geom *g, size_t nvol, uint64_t *cb, int always_sub);
void orc_fdk_cbct_roi(const orc_fdk_geom *g, size_t nvol, uint64_t *cb)
{
    orc_fdk_cbct_roi_impl(g, nvol, cb, 0);
}

/* FDKbp's own copy of the window (fdkbp.cpp:361-406) differs in one line:
 * "if (CBCT_roi[1]-8>=0) CBCT_roi[1]-=8" is always true for a size_t, so a
 * window starting at row 0 wraps to 2^64-8 (the single-precision module
 * tests CBCT_roi[1]!=0 first). */
void orc_fdkd_cbct_roi(const orc_fdk_geom *g, size_t nvol, uint64_t *cb)
{
    orc_fdk_cbct_roi_impl(g, nvol, cb, 1);
}

static void orc_fdk_cbct_roi_impl(const orc_fdk_geom *g, size_t nvol, uint64_t *cb, int always_sub)
{
    const float radius = (float)nvol * g->vox[0] / 2;
    const float pp1 = g->ppoint[1];
    const float r1 = (float)g->roi[1], r3 = (float)g->roi[3];
    const float vs = g->vox[0], sdd = g->sdd, sod = g->sod, res = g->pix[0];
    cb[0] = g->roi[0]; cb[2] = g->roi[2]; cb[1] = 0; cb[3] = 0;
    if (pp1 >= r1 && pp1 >= r3) {
        cb[3] = (uint64_t)(pp1 - ((pp1 - r3) * vs * sdd / (sod + radius)) / res);
        float value = pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res;
        if (value <= 0) cb[1] = 0;
        else cb[1] = (uint64_t)(pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res);
    }
    if (pp1 < r1 && pp1 < r3) {
        float value = pp1 + ((r1 - pp1) * vs * sdd / (sod + radius)) / res;
        cb[1] = (uint64_t)value;
        float value2 = pp1 + ((r3 - pp1) * vs * sdd / (sod - radius)) / res;
        if (value2 >= (float)g->proj_roi3) cb[3] = g->proj_roi3;
        else cb[3] = (uint64_t)value2;
    }
    if (pp1 >= r1 && pp1 < r3) {
        float value = pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res;
        if (value <= 0) cb[1] = 0;
        else cb[1] = (uint64_t)(pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res);
        float value2 = pp1 + ((r3 - pp1) * vs * sdd / (sod - radius)) / res;
        if (value2 >= (float)g->proj_roi3) cb[3] = g->proj_roi3;
        else cb[3] = (uint64_t)value2;
    }
    if (always_sub || cb[1] != 0) cb[1] -= 8;
    if (cb[3] + 8 <= g->proj_roi3) cb[3] += 8;
}

/* prepareFFT, fdkbp_single.cpp:960-963 */
size_t orc_fdk_padwidth(size_t width)
{
    return (size_t)(unsigned int)(2 * pow(2, ceil(log((double)(unsigned int)width) / log(2))));
}

/* ramp_filter_tuned, fdkbp_single.cpp:832-958, on one w x h image in place.
 * f64 != 0: evaluate in double (window and scale in double); the image stays
 * float only for the f32 flavour. */
static void orc_fdk_filter_impl(float *imgf, double *imgd, size_t w, size_t h, int kind)
{
    const unsigned int MARGINF = 5;
    const size_t P = orc_fdk_padwidth(w);
    const size_t padf = (P - w) / 2;
    const int f64 = imgd != NULL;
    double *pad = (double *)calloc(P * h, sizeof(double));
    for (size_t r = 0; r < h; ++r)
        for (size_t c = 0; c < w; ++c)
            pad[r * P + padf + c] = f64 ? imgd[r * w + c] : (double)imgf[r * w + c];
    for (size_t r = 0; r < MARGINF; ++r) memcpy(pad + r * P, pad + MARGINF * P, P * sizeof(double));
    for (size_t r = h - MARGINF; r < h; ++r)
        memcpy(pad + r * P, pad + (h - MARGINF - 1) * P, P * sizeof(double));
    for (size_t r = 0; r < h; ++r) {
        for (size_t c = 0; c < MARGINF; ++c) pad[r * P + c] = pad[r * P + MARGINF];
        for (size_t c = P - MARGINF; c < P; ++c) pad[r * P + c] = pad[r * P + P - MARGINF - 1];
    }
    float *rampf = (float *)malloc(sizeof(float) * P);
    double *rampd = (double *)malloc(sizeof(double) * P);
    const float DEGTORADF = ORC_FPI / 180.0f;
    for (size_t i = 0; i < P; ++i) {
        rampf[i] = (i < P / 2) ? (float)i : (float)(P - i);
        rampd[i] = rampf[i];
        rampf[i] = (float)(rampf[i] * (0.54 + 0.46 * (cosf((float)(unsigned)i * DEGTORADF * 360 / (float)(unsigned)P) + 1)));
        rampf[i] = (float)(rampf[i] / ((double)P * 0.5));
        if (kind == 0) {
            rampd[i] = rampd[i] * (0.54 + 0.46 * (cos(2.0 * ORC_DPI * (double)i / (double)P) + 1)) / ((double)P * 0.5);
        } else {
            /* extension: Shepp-Logan, |k| sinc(k / P) */
            const double k = rampd[i], x = ORC_DPI * k / (double)P;
            rampd[i] = k == 0.0 ? 0.0 : (k / ((double)P * 0.5)) * sin(x) / x;
        }
    }
    const float rpad = 1.0f / (float)P;
    orc_cplx *tw = orc_twiddles(P);
#pragma omp parallel
    {
        orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * P);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)h; ++r) {
            for (size_t i = 0; i < P; ++i) {
                a[i].re = f64 ? pad[r * P + i] : (double)(float)pad[r * P + i];
                a[i].im = 0.0;
            }
            orc_fft(a, P, -1, tw);
            for (size_t c = 0; c < P; ++c) {
                if (f64) { a[c].re *= rampd[c]; a[c].im *= rampd[c]; }
                else {
                    a[c].re = (double)((float)a[c].re * rampf[c]);
                    a[c].im = (double)((float)a[c].im * rampf[c]);
                }
            }
            orc_fft(a, P, +1, tw);
            for (size_t c = 0; c < w; ++c) {
                size_t i = padf + c;
                if (f64) imgd[r * w + c] = a[i].re / (double)P;
                else imgf[r * w + c] = (float)a[i].re * rpad;
            }
        }
        free(a);
    }
    free(tw); free(rampf); free(rampd); free(pad);
}

void orc_fdk_filter_f32(float *img, size_t w, size_t h) { orc_fdk_filter_impl(img, NULL, w, h, 0); }
void orc_fdk_filter_f64(double *img, size_t w, size_t h) { orc_fdk_filter_impl(NULL, img, w, h, 0); }
/* kind 0: the reference's ramp and roll-off; kind 1: the Shepp-Logan extension */
void orc_fdk_filter_ext_f64(double *img, size_t w, size_t h, int kind) { orc_fdk_filter_impl(NULL, img, w, h, kind); }

/* Full FDKbp_single::Process(3D) on one slab.
 * proj : np x h x w window projections (the rows the engine hands over),
 * vol  : nz x ny x nx (x fastest, MatrixXYZ) with nx = ny = roi[2]-roi[0] and
 *        nz = roi[3]-roi[1]; receives the mirrored/flipped final volume.
 * mfirst/msecond: circle mask over y (BuildCircleMask, XYZ variant). */
void orc_fdk_bp_f32(const float *proj, size_t w, size_t h, size_t np,
                    const float *angles, const float *weights, const orc_fdk_geom *g,
                    const size_t *mfirst, const size_t *msecond, float *vol)
{
    const size_t nx = (size_t)(g->roi[2] - g->roi[0]), ny = nx;
    const size_t nz = (size_t)(g->roi[3] - g->roi[1]);
    float *cv = (float *)calloc(nx * ny * nz, sizeof(float));
    float *img = (float *)malloc(sizeof(float) * w * h);
    const float scale = g->sod / g->sdd;
    const float sad_sid_2 = (g->sod * g->sod) / (g->sdd * g->sdd);
    const float sp[3] = {g->vox[0], g->vox[1], g->vox[2]};
    const float U = (float)(g->roi[2] - g->roi[0]);
    const float V = (float)(g->roi[3] - g->roi[1]);
    float origin[3];
    origin[0] = -(U - g->center) * sp[0] - sp[0] / 2;
    origin[1] = -(U - g->center) * sp[1] - sp[1] / 2;
    origin[2] = -(V - (g->ppoint[1] - (float)g->roi[1])) * sp[2] - sp[2] / 2;
    uint64_t cb[4];
    orc_fdk_cbct_roi(g, ny, cb);
    const float ic[2] = {g->ppoint[0] - (float)cb[0], g->ppoint[1] - (float)cb[1]};
    float *xip = (float *)malloc(sizeof(float) * 3 * nx);
    float *yip = (float *)malloc(sizeof(float) * 3 * ny);
    float *zip = (float *)malloc(sizeof(float) * 3 * nz);
    for (size_t ip = 0; ip < np; ++ip) {
        const float *src = proj + ip * w * h;
        for (size_t i = 0; i < w * h; ++i) img[i] = src[i] * weights[ip];
        float pm[12];
        orc_fdk_projmatrix_f32(g, angles[ip], pm);
        orc_fdk_filter_f32(img, w, h);
        for (size_t i = 0; i < w * h; ++i) { img[i] *= sad_sid_2; img[i] *= scale; }
        for (size_t i = 0; i < nx; ++i) {
            float x = (float)(origin[0] + (float)i * sp[0]);
            xip[i * 3 + 0] = x * (pm[0] + ic[0] * pm[8]);
            xip[i * 3 + 1] = x * (pm[4] + ic[1] * pm[8]);
            xip[i * 3 + 2] = x * pm[8];
        }
        for (size_t j = 0; j < ny; ++j) {
            float y = (float)(origin[1] + (float)j * sp[1]);
            yip[j * 3 + 0] = y * (pm[1] + ic[0] * pm[9]);
            yip[j * 3 + 1] = y * (pm[5] + ic[1] * pm[9]);
            yip[j * 3 + 2] = y * pm[9];
        }
        for (size_t k = 0; k < nz; ++k) {
            float z = (float)(origin[2] + (float)k * sp[2]);
            if (g->correct_tilt) {
                /* fdkbp_single.cpp:443-448: the centre of rotation of slice k
                 * (tan in double, the result held in a float) replaces pm[3] */
                const float pos = (float)cb[3] - (float)k - g->tilt_pivot;
                const float cor = (float)(tan((double)(-g->tilt_angle) * ORC_DPI / 180.0) * (double)pos +
                                          (double)g->center);
                pm[3] = ((cor - (g->ppoint[0] - (float)g->roi[0])) * g->vox[0]) / g->pix[0];
            }
            zip[k * 3 + 0] = z * (pm[2] + ic[0] * pm[10]) + ic[0] * pm[11] + pm[3];
            zip[k * 3 + 1] = z * (pm[6] + ic[1] * pm[10]) + ic[1] * pm[11] + pm[7];
            zip[k * 3 + 2] = z * pm[10] + pm[11];
        }
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)nz; ++k) {
            for (size_t j = 0; j < ny; ++j) {
                float acc2[3];
                acc2[0] = zip[3 * k] + yip[3 * j];
                acc2[1] = zip[3 * k + 1] + yip[3 * j + 1];
                acc2[2] = zip[3 * k + 2] + yip[3 * j + 2];
                float *row = cv + ((size_t)k * ny + j) * nx;
                for (size_t i = mfirst[j] + 1; i <= msecond[j]; ++i) {
                    float dw = 1.0f / (acc2[2] + xip[3 * i + 2]);
                    float r = dw * (acc2[1] + xip[3 * i + 1]);
                    float c = dw * (acc2[0] + xip[3 * i]);
                    float val = 0.0f;
                    r += 0.5f;
                    c += 0.5f;
                    if (!(r < 0) && !(r >= (float)h) && !(c < 0) && !(c >= (float)w))
                        val = img[(size_t)(int)r * w + (size_t)(int)c];
                    row[i] += dw * dw * val;
                }
            }
        }
    }
    /* mirror x, flip z: fdkreconbase.cpp:259-270 */
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                vol[((nz - 1 - k) * ny + j) * nx + i] = cv[(k * ny + j) * nx + (nx - 1 - i)];
    free(cv); free(img); free(xip); free(yip); free(zip);
}

static void orc_fdk_bp_f64_impl(const float *proj, size_t w, size_t h, size_t np,
                                const double *angles, const double *weights, const orc_fdk_geom *g,
                                const size_t *mfirst, const size_t *msecond, double *vol, int kind, int cosw);

/* FDK in double precision: geometry evaluated in double from the same
 * configuration, FDK filter in double, sampling/accumulation in double. */
void orc_fdk_bp_f64(const float *proj, size_t w, size_t h, size_t np,
                    const double *angles, const double *weights, const orc_fdk_geom *g,
                    const size_t *mfirst, const size_t *msecond, double *vol)
{
    orc_fdk_bp_f64_impl(proj, w, h, np, angles, weights, g, mfirst, msecond, vol, 0, 0);
}

/* The same with the project's filter extension (no reference counterpart):
 * kind 1 = Shepp-Logan ramp |k| sinc(k/P) / (P/2) instead of the reference's
 * window; cosw = 1 multiplies every detector pixel by the Feldkamp cosine
 * weight SDD / sqrt(SDD^2 + u^2 + v^2) (u, v in mm from the piercing point)
 * before the filter (and before the reference's edge-row copies). */
void orc_fdk_bp_ext_f64(const float *proj, size_t w, size_t h, size_t np,
                        const double *angles, const double *weights, const orc_fdk_geom *g,
                        const size_t *mfirst, const size_t *msecond, double *vol, int kind, int cosw)
{
    orc_fdk_bp_f64_impl(proj, w, h, np, angles, weights, g, mfirst, msecond, vol, kind, cosw);
}

static void orc_fdk_bp_f64_impl(const float *proj, size_t w, size_t h, size_t np,
                                const double *angles, const double *weights, const orc_fdk_geom *g,
                                const size_t *mfirst, const size_t *msecond, double *vol, int kind, int cosw)
{
    const size_t nx = (size_t)(g->roi[2] - g->roi[0]), ny = nx;
    const size_t nz = (size_t)(g->roi[3] - g->roi[1]);
    double *cv = (double *)calloc(nx * ny * nz, sizeof(double));
    double *img = (double *)malloc(sizeof(double) * w * h);
    const double sod = g->sod, sdd = g->sdd;
    const double scale = sod / sdd, sad_sid_2 = (sod * sod) / (sdd * sdd);
    const double sp[3] = {g->vox[0], g->vox[1], g->vox[2]};
    const double U = (double)(g->roi[2] - g->roi[0]);
    const double V = (double)(g->roi[3] - g->roi[1]);
    double origin[3];
    origin[0] = -(U - g->center) * sp[0] - sp[0] / 2;
    origin[1] = -(U - g->center) * sp[1] - sp[1] / 2;
    origin[2] = -(V - ((double)g->ppoint[1] - (double)g->roi[1])) * sp[2] - sp[2] / 2;
    uint64_t cb[4];
    orc_fdk_cbct_roi(g, ny, cb);
    const double ic[2] = {(double)g->ppoint[0] - (double)cb[0], (double)g->ppoint[1] - (double)cb[1]};
    const double offCenter = ((double)g->ppoint[0] - (double)g->roi[0] - g->center) * sp[0];
    for (size_t ip = 0; ip < np; ++ip) {
        const float *src = proj + ip * w * h;
        for (size_t i = 0; i < w * h; ++i) img[i] = (double)src[i] * weights[ip];
        if (cosw)
            for (size_t r = 0; r < h; ++r)
                for (size_t c = 0; c < w; ++c) {
                    const double u = ((double)c - ic[0]) * g->pix[0], v = ((double)r - ic[1]) * g->pix[1];
                    img[r * w + c] *= sdd / sqrt(sdd * sdd + u * u + v * v);
                }
        orc_fdk_filter_impl(NULL, img, w, h, kind);
        for (size_t i = 0; i < w * h; ++i) img[i] *= sad_sid_2 * scale;
        /* projection matrix in closed form: rows -plt/pix_u, -pup/pix_v, -nrm/sdd */
        double a = (g->direction == 0 ? -angles[ip] : angles[ip]) * ORC_DPI / 180.0;
        double nr[3] = {cos(a), sin(a), 0.0};
        double pl[3] = {nr[1], -nr[0], 0.0};
        double pu[3] = {pl[1] * nr[2] - pl[2] * nr[1], pl[2] * nr[0] - pl[0] * nr[2], pl[0] * nr[1] - pl[1] * nr[0]};
        double pm[12] = {-pl[0] / g->pix[0], -pl[1] / g->pix[0], -pl[2] / g->pix[0], -offCenter / g->pix[0],
                         -pu[0] / g->pix[1], -pu[1] / g->pix[1], -pu[2] / g->pix[1], 0.0,
                         -nr[0] / sdd, -nr[1] / sdd, -nr[2] / sdd, sod / sdd};
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)nz; ++k) {
            double z = origin[2] + (double)k * sp[2];
            for (size_t j = 0; j < ny; ++j) {
                double y = origin[1] + (double)j * sp[1];
                double *row = cv + ((size_t)k * ny + j) * nx;
                for (size_t i = mfirst[j] + 1; i <= msecond[j]; ++i) {
                    double x = origin[0] + (double)i * sp[0];
                    double p0 = pm[0] * x + pm[1] * y + pm[2] * z + pm[3];
                    double p1 = pm[4] * x + pm[5] * y + pm[6] * z + pm[7];
                    double p2 = pm[8] * x + pm[9] * y + pm[10] * z + pm[11];
                    double dw = 1.0 / p2;
                    double c = dw * (p0 + ic[0] * p2) + 0.5;
                    double r = dw * (p1 + ic[1] * p2) + 0.5;
                    double val = 0.0;
                    if (!(r < 0) && !(r >= (double)h) && !(c < 0) && !(c >= (double)w))
                        val = img[(size_t)(int)r * w + (size_t)(int)c];
                    row[i] += dw * dw * val;
                }
            }
        }
    }
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                vol[((nz - 1 - k) * ny + j) * nx + i] = cv[(k * ny + j) * nx + (nx - 1 - i)];
    free(cv); free(img);
}

/* ------------------------------------------------------------------------ */
/* FDKbp (double-precision geometry), FDK/fdkbp.cpp                         */
/* ------------------------------------------------------------------------ */

/* getProjMatrix, fdkbp.cpp:83-257: camera, panel vectors and both matrices in
 * double; the product accumulates in a float ("float acc", :259-266), so
 * every matrix entry is a float value held in a double.  offCenter and
 * 1/fResolution are float expressions (float operands). */
void orc_fdkd_projmatrix(const orc_fdk_geom *g, float angle, double *pm /*12*/)
{
    double ext[16] = {0}, intr[12] = {0};
    const double sad = g->sod, sid = g->sdd;
    const double offCenter = (double)(((g->ppoint[0] - (float)g->roi[0]) - g->center) * g->vox[0]);
    double cam[3] = {0.0, 0.0, 0.0};
    if (g->direction == 0) {
        cam[0] = sad * cos(-angle * ORC_DPI / 180);
        cam[1] = sad * sin(-angle * ORC_DPI / 180);
    } else {
        cam[0] = sad * cos(angle * ORC_DPI / 180);
        cam[1] = sad * sin(angle * ORC_DPI / 180);
    }
    const double vup[3] = {0.0, 0.0, 1.0}, tgt[3] = {0.0, 0.0, 0.0};
    double nrm[3], plt[3], pup[3];
    nrm[0] = cam[0] - tgt[0]; nrm[1] = cam[1] - tgt[1]; nrm[2] = cam[2] - tgt[2];
    double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    nrm[0] /= nn; nrm[1] /= nn; nrm[2] /= nn;
    plt[0] = nrm[1] * vup[2] - nrm[2] * vup[1];
    plt[1] = nrm[2] * vup[0] - nrm[0] * vup[2];
    plt[2] = nrm[0] * vup[1] - nrm[1] * vup[0];
    double np_ = sqrt(plt[0] * plt[0] + plt[1] * plt[1] + plt[2] * plt[2]);
    plt[0] /= np_; plt[1] /= np_; plt[2] /= np_;
    pup[0] = plt[1] * nrm[2] - plt[2] * nrm[1];
    pup[1] = plt[2] * nrm[0] - plt[0] * nrm[2];
    pup[2] = plt[0] * nrm[1] - plt[1] * nrm[0];
    double nu_ = sqrt(pup[0] * pup[0] + pup[1] * pup[1] + pup[2] * pup[2]);
    pup[0] /= nu_; pup[1] /= nu_; pup[2] /= nu_;
    ext[0] = -1.0 * plt[0]; ext[1] = -1.0 * plt[1]; ext[2] = -1.0 * plt[2];
    ext[4] = -1.0 * pup[0]; ext[5] = -1.0 * pup[1]; ext[6] = -1.0 * pup[2];
    ext[8] = -1.0 * nrm[0]; ext[9] = -1.0 * nrm[1]; ext[10] = -1.0 * nrm[2];
    ext[15] = 1.0;
    ext[3] = plt[0] * tgt[0] + plt[1] * tgt[1] + plt[2] * tgt[2] - offCenter;
    ext[7] = pup[0] * tgt[0] + pup[1] * tgt[1] + pup[2] * tgt[2];
    ext[11] = nrm[0] * tgt[0] + nrm[1] * tgt[1] + nrm[2] * tgt[2] + sad;
    intr[0] = (double)(1 / g->pix[0]);
    intr[5] = (double)(1 / g->pix[1]);
    intr[10] = 1 / sid;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc = (float)((double)acc + intr[i * 4 + k] * ext[k * 4 + j]);
            pm[i * 4 + j] = acc;
        }
}

/* FDKbp::ramp_filter_tuned, fdkbp.cpp:822-948: the float image padded as in
 * FDKbp_single, transformed in double (FFTW double c2c there, the exact radix-2
 * DFT here), ramp |k| x (0.54+0.46(cos(i*DEGTORAD*360/P)+1)) / (P*0.5) in
 * double, inverse, /P in double, rounded to float. */
void orc_fdkd_filter(float *img, size_t w, size_t h)
{
    const unsigned int MARGIN = 5;
    const size_t P = orc_fdk_padwidth(w);
    const size_t padf = (P - w) / 2;
    float *pad = (float *)calloc(P * h, sizeof(float));
    for (size_t r = 0; r < h; ++r) memcpy(pad + r * P + padf, img + r * w, w * sizeof(float));
    for (size_t r = 0; r < MARGIN; ++r) memcpy(pad + r * P, pad + MARGIN * P, P * sizeof(float));
    for (size_t r = h - MARGIN; r < h; ++r) memcpy(pad + r * P, pad + (h - MARGIN - 1) * P, P * sizeof(float));
    for (size_t r = 0; r < h; ++r) {
        for (size_t c = 0; c < MARGIN; ++c) pad[r * P + c] = pad[r * P + MARGIN];
        for (size_t c = P - MARGIN; c < P; ++c) pad[r * P + c] = pad[r * P + P - MARGIN - 1];
    }
    double *ramp = (double *)malloc(sizeof(double) * P);
    const double DEGTORAD = ORC_DPI / 180.0;
    for (size_t i = 0; i < P / 2; ++i) ramp[i] = (double)i;
    for (size_t i = P / 2; i < P; ++i) ramp[i] = (double)(P - i);
    for (size_t i = 0; i < P; ++i) {
        ramp[i] *= (0.54 + 0.46 * (cos((double)(unsigned)i * DEGTORAD * 360 / (double)(unsigned)P) + 1));
        ramp[i] /= (double)(unsigned)P * 0.5;
    }
    orc_cplx *tw = orc_twiddles(P);
#pragma omp parallel
    {
        orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * P);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)h; ++r) {
            for (size_t i = 0; i < P; ++i) { a[i].re = (double)pad[r * P + i]; a[i].im = 0.0; }
            orc_fft(a, P, -1, tw);
            for (size_t c = 0; c < P; ++c) { a[c].re *= ramp[c]; a[c].im *= ramp[c]; }
            orc_fft(a, P, +1, tw);
            for (size_t c = 0; c < w; ++c) img[r * w + c] = (float)(a[padf + c].re / (double)(unsigned)P);
        }
        free(a);
    }
    free(tw); free(ramp); free(pad);
}

/* Full FDKbp::Process(3D) on one slab (fdkreconbase.cpp:217-277 with
 * fdkbp.cpp:272-510, 612-631).  Differences from FDKbp_single:
 *   - scale = SOD/SDD/nProj (float), nProj = projections in the call;
 *   - img *= sad_sid_2 with sad_sid_2 a double holding the float
 *     (SOD*SOD)/(SDD*SDD); img *= scale in float;
 *   - ic, origin, spacing as float expressions; xip/yip/zip, the depth, its
 *     reciprocal and the sampling positions in double;
 *   - get_pixel_value_c takes double r, c (r += 0.5, bounds, truncation);
 *   - volume (float) += dw*dw*val evaluated in double;
 *   - the window origin of ic[1] uses orc_fdkd_cbct_roi (see there). */
void orc_fdkd_bp(const float *proj, size_t w, size_t h, size_t np,
                 const float *angles, const float *weights, const orc_fdk_geom *g,
                 const size_t *mfirst, const size_t *msecond, float *vol)
{
    const size_t nx = (size_t)(g->roi[2] - g->roi[0]), ny = nx;
    const size_t nz = (size_t)(g->roi[3] - g->roi[1]);
    float *cv = (float *)calloc(nx * ny * nz, sizeof(float));
    float *img = (float *)malloc(sizeof(float) * w * h);
    const float scale = g->sod / g->sdd / (float)np;
    const double sad_sid_2 = (double)((g->sod * g->sod) / (g->sdd * g->sdd));
    const float sp[3] = {g->vox[0], g->vox[1], g->vox[2]};
    const float U = (float)(g->roi[2] - g->roi[0]);
    const float V = (float)(g->roi[3] - g->roi[1]);
    float origin[3];
    origin[0] = -(U - g->center) * sp[0] - sp[0] / 2;
    origin[1] = -(U - g->center) * sp[1] - sp[1] / 2;
    origin[2] = -(V - (g->ppoint[1] - (float)g->roi[1])) * sp[2] - sp[2] / 2;
    uint64_t cb[4];
    orc_fdkd_cbct_roi(g, ny, cb);
    const double ic[2] = {(double)(g->ppoint[0] - (float)cb[0]), (double)(g->ppoint[1] - (float)cb[1])};
    double *xip = (double *)malloc(sizeof(double) * 3 * nx);
    double *yip = (double *)malloc(sizeof(double) * 3 * ny);
    double *zip = (double *)malloc(sizeof(double) * 3 * nz);
    for (size_t ip = 0; ip < np; ++ip) {
        const float *src = proj + ip * w * h;
        for (size_t i = 0; i < w * h; ++i) img[i] = src[i] * weights[ip];
        double pm[12];
        orc_fdkd_projmatrix(g, angles[ip], pm);
        orc_fdkd_filter(img, w, h);
        for (size_t i = 0; i < w * h; ++i) {
            img[i] = (float)((double)img[i] * sad_sid_2);
            img[i] *= scale;
        }
        for (size_t i = 0; i < nx; ++i) {
            double x = (double)(origin[0] + (float)(long)i * sp[0]);
            xip[i * 3 + 0] = x * (pm[0] + ic[0] * pm[8]);
            xip[i * 3 + 1] = x * (pm[4] + ic[1] * pm[8]);
            xip[i * 3 + 2] = x * pm[8];
        }
        for (size_t j = 0; j < ny; ++j) {
            double y = (double)(origin[1] + (float)(long)j * sp[1]);
            yip[j * 3 + 0] = y * (pm[1] + ic[0] * pm[9]);
            yip[j * 3 + 1] = y * (pm[5] + ic[1] * pm[9]);
            yip[j * 3 + 2] = y * pm[9];
        }
        for (size_t k = 0; k < nz; ++k) {
            double z = (double)(origin[2] + (float)(long)k * sp[2]);
            if (g->correct_tilt) {
                /* fdkbp.cpp:452-458, in double, with FDKbp's own window row cb[3] */
                const double pos = (double)cb[3] - (double)k - (double)g->tilt_pivot;
                const double cor = tan((double)(-g->tilt_angle) * ORC_DPI / 180.0) * pos + (double)g->center;
                pm[3] = ((cor - (double)(g->ppoint[0] - (float)g->roi[0])) * (double)g->vox[0]) / (double)g->pix[0];
            }
            zip[k * 3 + 0] = z * (pm[2] + ic[0] * pm[10]) + ic[0] * pm[11] + pm[3];
            zip[k * 3 + 1] = z * (pm[6] + ic[1] * pm[10]) + ic[1] * pm[11] + pm[7];
            zip[k * 3 + 2] = z * pm[10] + pm[11];
        }
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)nz; ++k) {
            for (size_t j = 0; j < ny; ++j) {
                double acc2[3];
                acc2[0] = zip[3 * k] + yip[3 * j];
                acc2[1] = zip[3 * k + 1] + yip[3 * j + 1];
                acc2[2] = zip[3 * k + 2] + yip[3 * j + 2];
                float *row = cv + ((size_t)k * ny + j) * nx;
                for (size_t i = mfirst[j] + 1; i <= msecond[j]; ++i) {
                    double acc3[3];
                    acc3[0] = acc2[0] + xip[3 * i];
                    acc3[1] = acc2[1] + xip[3 * i + 1];
                    acc3[2] = acc2[2] + xip[3 * i + 2];
                    const double dw = 1.0 / acc3[2];
                    double r = acc3[1] * dw, c = acc3[0] * dw;
                    float val = 0.0f;
                    r += 0.5;
                    c += 0.5;
                    if (!(r < 0) && !(r >= (double)h) && !(c < 0) && !(c >= (double)w))
                        val = img[(size_t)(int)r * w + (size_t)(int)c];
                    row[i] = (float)((double)row[i] + dw * dw * (double)val);
                }
            }
        }
    }
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                vol[((nz - 1 - k) * ny + j) * nx + i] = cv[(k * ny + j) * nx + (nx - 1 - i)];
    free(cv); free(img); free(xip); free(yip); free(zip);
}

/* ------------------------------------------------------------------------ */
/* FullLogNorm (StdPreprocModules/src/NormPlugins.cpp), non-LUT path.  The   */
/* reference's log(float) resolves to the float overload (sharedbuffer.h     */
/* pulls <cmath> and "using namespace std" into every kipl image user).      */
/* ------------------------------------------------------------------------ */

/* FullLogNorm::SetReferenceImages (:300-341): the flat field becomes
 * log((flat - dark) / flat_dose), 0 where flat - dark <= 0; dark == NULL when
 * no dark current images are configured (nDCCount == 0). */
void orc_lognorm_flat(float *flat, const float *dark, size_t n, float flat_dose)
{
    const float dose = 1.0f / flat_dose;
    for (size_t i = 0; i < n; ++i) {
        const float p = dark ? flat[i] - dark[i] : flat[i];
        flat[i] = p <= 0 ? 0.0f : logf(p * dose);
    }
}

/* FullLogNorm::ProcessCore(3D) (:411-558, bUseLUT == false) on n_proj
 * projections of n_pix pixels.  flat_log: the prepared flat (NULL when
 * nOBCount == 0); dark: NULL when nDCCount == 0; dose: per-projection doses
 * (NULL when usenormregion is false, i.e. one dose log of 0). */
void orc_lognorm_f32(float *img, size_t n_proj, size_t n_pix, const float *flat_log, const float *dark,
                     const float *dose, float dark_dose)
{
#pragma omp parallel for schedule(static)
    for (long j = 0; j < (long)n_proj; ++j) {
        float dl = 0.0f;
        if (dose) {
            const float d = dose[j] - dark_dose;
            dl = logf(d < 1 ? 1.0f : d);
        }
        float *p = img + (size_t)j * n_pix;
        for (size_t i = 0; i < n_pix; ++i) {
            const float v = dark ? p[i] - dark[i] : p[i];
            if (v <= 0.0f) p[i] = 0.0f;
            else p[i] = flat_log ? flat_log[i] - logf(v) + dl : -logf(v) + dl;
        }
    }
}

/* FullLogNorm's look-up-table path (bUseLUT).  The module's LUT is kipl's
 * LogLUT(1<<16, 0, 65535) (NormPlugins.cpp:245, LUTCollection.cpp:11-36,
 * LUTbase.cpp:17-24): slope (65536-1)/65535 = 1, intercept 0.5, entry 0 =
 * log(1/slope) = 0, entry i = log(i) (x stepping from 1 by 1, exact in float;
 * log of a float there takes the double overload, rounded to float).
 * LUTbase::BasicLUT (:60-78): v < 0 -> entry 0, 65535 < v -> entry 65535,
 * else entry (size_t)(v * 1 + 0.5). */
void orc_loglut_build(float *lut /* 65536 */)
{
    lut[0] = (float)log(1.0);
    for (size_t i = 1; i < 65536; ++i) lut[i] = (float)log((double)(float)i);
}

static float orc_loglut(const float *lut, float v)
{
    if (v < 0.0f) return lut[0];
    if (65535.0f < v) return lut[65535];
    return lut[(size_t)(v * 1.0f + 0.5f)];
}

/* ProcessCore(3D) with bUseLUT (NormPlugins.cpp:455-475) and ProcessCore(2D)
 * (:370-381): p = img - dark (nDCCount != 0) or img, L = LUT(p) (no p <= 0
 * test), img = flat_log - L + term[j], or -L + term[j] without a flat.
 * term[j]: 3-D with open beams doselist[j]; 3-D without doselist[j] -
 * doselist[0]; 2-D cLogDose (orc_lognorm_lut_terms makes the 3-D ones). */
void orc_lognorm_lut_f32(float *img, size_t n_proj, size_t n_pix, const float *flat_log, const float *dark,
                         const float *term)
{
    float *lut = (float *)malloc(sizeof(float) * 65536);
    orc_loglut_build(lut);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < (long)n_proj; ++j) {
        const float t = term ? term[j] : 0.0f;
        float *p = img + (size_t)j * n_pix;
        for (size_t i = 0; i < n_pix; ++i) {
            const float v = dark ? p[i] - dark[i] : p[i];
            const float L = orc_loglut(lut, v);
            p[i] = flat_log ? flat_log[i] - L + t : -L + t;
        }
    }
    free(lut);
}

/* The 3-D path's per-projection terms (NormPlugins.cpp:418-428, 459, 470-471):
 * doselist[j] = log(max(dose[j] - dark_dose, 1)) with usenormregion, else 0;
 * without open beams the term is doselist[j] - doselist[0]. */
void orc_lognorm_lut_terms(const float *dose, size_t n_proj, float dark_dose, int has_flat, float *term)
{
    for (size_t j = 0; j < n_proj; ++j) {
        float dl = 0.0f;
        if (dose) {
            const float d = dose[j] - dark_dose;
            dl = logf(d < 1 ? 1.0f : d);
        }
        term[j] = dl;
    }
    if (!has_flat) {
        const float d0 = term[0];
        for (size_t j = 0; j < n_proj; ++j) term[j] = term[j] - d0;
    }
}

/* ------------------------------------------------------------------------ */
/* Matched Joseph projector pair in double precision (the project's own     */
/* operator; no reference counterpart).  Geometry of the parallel-beam BP:  */
/*   u(x,y) = center + cos(phi)(y+1-c) - sin(phi)(x+1-c), c = nx/2,          */
/* kernel K(t) = max(0, 1-|t|/m)/m with m = max(|sin|,|cos|).               */
/* Forward is ray-driven (Joseph stepping), adjoint voxel-driven.           */
/* vol [nx(y)][nx(x)][nz], sino [na][nd][nz].                               */
/* ------------------------------------------------------------------------ */
static void orc_jtab(double ang, int direction, size_t nx, double center,
                     double *su, double *c, double *s, double *m)
{
    double dirw = direction == 0 ? -1.0 : 1.0;
    double phi = dirw * ang * ORC_DPI / 180.0;
    *s = sin(phi); *c = cos(phi);
    *su = (double)(nx / 2) * (*s - *c) + center;
    *m = fabs(*s) > fabs(*c) ? fabs(*s) : fabs(*c);
}

void orc_joseph_forward_f64(const double *vol, size_t nx, size_t nz, size_t nd, double center,
                            int direction, const double *angles, size_t na, double *sino)
{
    const long N = (long)nx;
#pragma omp parallel for schedule(dynamic, 1)
    for (long ia = 0; ia < (long)na; ++ia) {
        double su, c, s, m;
        orc_jtab(angles[ia], direction, nx, center, &su, &c, &s, &m);
        for (size_t n = 0; n < nd; ++n) {
            double *out = sino + ((size_t)ia * nd + n) * nz;
            for (size_t z = 0; z < nz; ++z) out[z] = 0.0;
            for (long k = 0; k < N; ++k) {
                double pos = fabs(s) >= fabs(c) ? (su + c * (double)(k + 1) - (double)n) / s - 1.0
                                                : ((double)n - su + s * (double)(k + 1)) / c - 1.0;
                double fl = floor(pos), f = pos - fl;
                long j0 = (long)fl;
                for (int t = 0; t < 2; ++t) {
                    long j = j0 + t;
                    double w = t == 0 ? 1.0 - f : f;
                    if (j < 0 || j >= N || w == 0.0) continue;
                    const double *v = fabs(s) >= fabs(c) ? vol + ((size_t)k * nx + (size_t)j) * nz
                                                          : vol + ((size_t)j * nx + (size_t)k) * nz;
                    for (size_t z = 0; z < nz; ++z) out[z] += w * v[z];
                }
            }
            for (size_t z = 0; z < nz; ++z) out[z] /= m;
        }
    }
}

void orc_joseph_adjoint_f64(const double *sino, size_t nx, size_t nz, size_t nd, double center,
                            int direction, const double *angles, size_t na, double *vol)
{
    memset(vol, 0, sizeof(double) * nx * nx * nz);
    double *tab = (double *)malloc(sizeof(double) * 4 * na);
    for (size_t ia = 0; ia < na; ++ia)
        orc_jtab(angles[ia], direction, nx, center, &tab[4 * ia], &tab[4 * ia + 1], &tab[4 * ia + 2], &tab[4 * ia + 3]);
#pragma omp parallel for schedule(static)
    for (long y = 0; y < (long)nx; ++y)
        for (size_t x = 0; x < nx; ++x) {
            double *out = vol + ((size_t)y * nx + x) * nz;
            for (size_t ia = 0; ia < na; ++ia) {
                double su = tab[4 * ia], c = tab[4 * ia + 1], s = tab[4 * ia + 2], m = tab[4 * ia + 3];
                double u = su + c * (double)(y + 1) - s * (double)(x + 1);
                long n0 = (long)floor(u);
                for (long n = n0; n <= n0 + 1; ++n) {
                    if (n < 0 || n >= (long)nd) continue;
                    double w = 1.0 - fabs(u - (double)n) / m;
                    if (w <= 0.0) continue;
                    w /= m;
                    const double *p = sino + ((size_t)ia * nd + (size_t)n) * nz;
                    for (size_t z = 0; z < nz; ++z) out[z] += w * p[z];
                }
            }
        }
    free(tab);
}

/* ------------------------------------------------------------------------
 * LinearForwardProjector::project (2-D)
 * (frameworks/tomography/Framework/ReconAlgorithms/ReconAlgorithms/
 * linearforwardprojector.cpp:15-143) with the reference's types: float
 * state, double dX*sinT/cosT and scale factors, the sequential float
 * position update nfloat += alpha, truncation, the [0, N-1) validity test and
 * the ceil/floor step range with eps = 1e-4.  slice(x, y) = slice[x + y*M];
 * proj(r, t) = proj[r + t*M] (R = M).  The reference's trig is std::sin /
 * std::cos of a float (sharedbuffer.h brings `using namespace std`).
 * ------------------------------------------------------------------------ */
void orc_linear_forward_f32(const float *slice, size_t Ms, const float *angles, size_t T, float *proj)
{
    const int M = (int)Ms, N = (int)Ms, R = (int)Ms;
    const float eps = 1e-4f;
    const float rho_min = (float)(-R / 2), x_min = (float)(-M / 2);
    const float Delta_rho = 1.0f, Delta_x = 1.0f;
    memset(proj, 0, sizeof(float) * (size_t)R * T);
    for (size_t t = 0; t < T; ++t) {
        float theta = angles[t] + 0.0f;
        theta *= ORC_FPI / 180.0f;
        const float sintheta = sinf(theta), costheta = cosf(theta);
        const float rhooffset = x_min * (costheta + sintheta);
        if ((double)fabsf(sintheta) > sqrt(0.5)) {
            const float alpha = -costheta / sintheta;
            const double dXsinT = (double)(Delta_x * sintheta);
            const double dXdsinT = (double)(Delta_x / fabsf(sintheta));
            float rho = 0.0f;
            for (int r = 0; r < R; ++r, rho += Delta_rho) {
                const float beta = (float)((double)(rho + rho_min - rhooffset) / dXsinT);
                const float betap = (float)((double)beta + 0.5);
                float sum = 0.0f;
                int mmin, mmax;
                if (alpha > 1e-6f) {
                    mmin = (int)ceilf(-(betap - eps) / alpha);
                    mmax = 1 + (int)floorf(((float)N - betap - eps) / alpha);
                } else if (alpha < -1e-6f) {
                    mmin = (int)ceilf(((float)N - betap - eps) / alpha);
                    mmax = 1 + (int)floorf(-(betap - eps) / alpha);
                } else {
                    mmin = 0;
                    mmax = M;
                }
                if (mmin < 0) mmin = 0;
                if (mmin >= M) mmin = M - 1;
                if (mmax > M) mmax = M;
                float nfloat = beta + (float)mmin * alpha;
                for (int m = mmin; m < mmax; ++m) {
                    const float nf = (float)(int)nfloat;
                    const float w = nfloat - nf;
                    if (nfloat >= 0.0f && nfloat < (float)(N - 1)) {
                        const float a = slice[(size_t)m + (size_t)nf * M];
                        const float b = slice[(size_t)m + (size_t)(nf + 1.0f) * M];
                        sum += a + w * (b - a);
                    }
                    nfloat += alpha;
                }
                proj[(size_t)r + t * R] = (float)((double)sum * dXdsinT);
            }
        } else {
            const float alpha = -sintheta / costheta;
            const double dXcosT = (double)(Delta_x * costheta);
            const double dXdcosT = (double)(Delta_x / fabsf(costheta));
            float rho = 0.0f;
            for (int r = 0; r < R; ++r, rho += Delta_rho) {
                const float beta = (float)((double)(rho + rho_min - rhooffset) / dXcosT);
                const float betap = (float)((double)beta + 0.5);
                float sum = 0.0f;
                int nmin, nmax;
                if (alpha > 1e-6f) {
                    nmin = (int)ceilf(-(betap - eps) / alpha);
                    nmax = 1 + (int)floorf(((float)M - betap - eps) / alpha);
                } else if (alpha < -1e-6f) {
                    nmin = (int)ceilf(((float)M - betap - eps) / alpha);
                    nmax = 1 + (int)floorf(-(betap - eps) / alpha);
                } else {
                    nmin = 0;
                    nmax = N;
                }
                if (nmin < 0) nmin = 0;
                if (nmin >= N) nmin = N - 1;
                if (nmax > N) nmax = N;
                float mfloat = beta + (float)nmin * alpha;
                for (int n = nmin; n < nmax; ++n) {
                    const float mf = (float)(int)mfloat;
                    const float w = mfloat - mf;
                    if (mfloat >= 0.0f && mfloat < (float)(M - 1)) {
                        const float a = slice[(size_t)mf + (size_t)n * M];
                        const float b = slice[(size_t)(mf + 1.0f) + (size_t)n * M];
                        sum += a + w * (b - a);
                    }
                    mfloat += alpha;
                }
                proj[(size_t)r + t * R] = (float)((double)sum * dXdcosT);
            }
        }
    }
}
