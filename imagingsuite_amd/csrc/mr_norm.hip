// FullLogNorm on gfx950 (frameworks/tomography/Preprocessing/StdPreprocModules/
// src/NormPlugins.cpp, FullLogNorm::SetReferenceImages :300-341 and
// ProcessCore(2D/3D) :343-558, the non-LUT path and the LUT path).
//
//   reference:  flat_log = log((flat - dark) / flat_dose), 0 where <= 0
//   projection: p = img - dark;  img = p <= 0 ? 0 : flat_log - log(p) + dose_log[j]
//               (no flat: -log(p) + dose_log[j]; no dark: p = img)
//   dose_log[j] = log(max(dose[j] - dark_dose, 1)) per projection, evaluated
//               on the host like the reference (one value per projection).
//
// Element-wise and HBM-bound (8 bytes per pixel without references, up to 16
// with flat and dark, which stay in L2 across projections).  Each thread
// handles float4 groups of one projection row range; log is the device
// logf (within 1 ulp of the host's).
#include "mr_common.h"
#include "mr_filter.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace mr {

// FullLogNorm's look-up-table path (bUseLUT): kipl's LogLUT(1<<16, 0, 65535)
// (NormPlugins.cpp:245; LUTCollection.cpp:11-36, LUTbase.cpp:17-24) has
// slope 1 and intercept 0.5; entry 0 is log(1) and entry i is log(i), the
// double log rounded to float.  The table is computed on the host and kept in
// one array per GPU, written once per device before its first use.
constexpr int kLogLut = 65536;
__device__ __align__(256) float g_log_lut[kLogLut];

// LUTbase::BasicLUT (LUTbase.cpp:60-78): below 0 entry 0, above 65535 the
// last entry, else entry (size_t)(v * slope + 0.5) with slope 1.
__device__ __forceinline__ float loglut_px(float v, const float *flat_log, const float *dark, size_t i, float t,
                                           const float *lut)
{
    const float p = dark ? v - dark[i] : v;
    float L;
    if (p < 0.0f) L = lut[0];
    else if (65535.0f < p) L = lut[kLogLut - 1];
    else L = lut[(unsigned)(p + 0.5f)];
    return flat_log ? flat_log[i] - L + t : -L + t;
}

// blockIdx.y = projection; t: per-projection additive term (NULL: 0).
__global__ __launch_bounds__(256) void lognorm_lut_kernel(float *img, size_t n_pix, const float *flat_log,
                                                          const float *dark, const float *term, const float *lut)
{
    const size_t j = blockIdx.y;
    const float t = term ? term[j] : 0.0f;
    float *p = img + j * n_pix;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if ((n_pix & 3) == 0) {
        float4 *p4 = reinterpret_cast<float4 *>(p);
        for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_pix / 4; q += stride) {
            float4 v = p4[q];
            v.x = loglut_px(v.x, flat_log, dark, 4 * q + 0, t, lut);
            v.y = loglut_px(v.y, flat_log, dark, 4 * q + 1, t, lut);
            v.z = loglut_px(v.z, flat_log, dark, 4 * q + 2, t, lut);
            v.w = loglut_px(v.w, flat_log, dark, 4 * q + 3, t, lut);
            p4[q] = v;
        }
    } else {
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_pix; i += stride)
            p[i] = loglut_px(p[i], flat_log, dark, i, t, lut);
    }
}

// Device address of this GPU's table, filled on the first call per device.
static int log_lut_device(const float **out)
{
    static std::mutex mu;
    static std::map<int, const float *> ready;
    int dev = 0;
    MR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ready.find(dev);
    if (it != ready.end()) {
        *out = it->second;
        return MR_OK;
    }
    std::vector<float> h(kLogLut);
    h[0] = (float)std::log(1.0);
    for (int i = 1; i < kLogLut; ++i) h[i] = (float)std::log((double)(float)i);
    void *ptr = nullptr;
    MR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_log_lut)));
    MR_HIP(hipMemcpy(ptr, h.data(), sizeof(float) * kLogLut, hipMemcpyHostToDevice));
    ready[dev] = static_cast<const float *>(ptr);
    *out = static_cast<const float *>(ptr);
    return MR_OK;
}

__global__ __launch_bounds__(256) void lognorm_reference_kernel(float *flat, const float *dark, size_t n, float dose)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float p = dark ? flat[i] - dark[i] : flat[i];
        flat[i] = p <= 0.0f ? 0.0f : logf(p * dose);
    }
}

// blockIdx.y = projection; threads stride over its pixels.
__global__ __launch_bounds__(256) void lognorm_kernel(float *img, size_t n_pix, const float *flat_log,
                                                      const float *dark, const float *dose_log)
{
    const size_t j = blockIdx.y;
    const float dl = dose_log ? dose_log[j] : 0.0f;
    float *p = img + j * n_pix;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if ((n_pix & 3) == 0) {
        float4 *p4 = reinterpret_cast<float4 *>(p);
        for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_pix / 4; q += stride) {
            float4 v = p4[q];
            v.x = lognorm_px(v.x, flat_log, dark, 4 * q + 0, dl);
            v.y = lognorm_px(v.y, flat_log, dark, 4 * q + 1, dl);
            v.z = lognorm_px(v.z, flat_log, dark, 4 * q + 2, dl);
            v.w = lognorm_px(v.w, flat_log, dark, 4 * q + 3, dl);
            p4[q] = v;
        }
    } else {
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_pix; i += stride)
            p[i] = lognorm_px(p[i], flat_log, dark, i, dl);
    }
}

}  // namespace mr

extern "C" int mr_lognorm_reference(float *d_flat, const float *d_dark, size_t n_pix, float flat_dose, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(d_flat != nullptr && n_pix > 0, "NULL flat field or empty image");
    const float dose = 1.0f / flat_dose;
    MR_REQUIRE(dose == dose, "The reference dose is a NaN");
    MR_REQUIRE(!std::isinf(dose), "The reference dose is Inf");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = (unsigned)std::min<size_t>((n_pix + 255) / 256, 65536);
    mr::ProfScope prof("lognorm_reference", s);
    hipLaunchKernelGGL(mr::lognorm_reference_kernel, dim3(blocks), dim3(256), 0, s, d_flat, d_dark, n_pix, dose);
    MR_HIP(hipGetLastError());
    return MR_OK;
    MR_GUARD_END
}

namespace {

// dl: host table of n_proj dose logs, or NULL for a dose term of 0.
int run_lognorm(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                const float *dl, hipStream_t s)
{
    float *d_dl = nullptr;
    if (dl) {
        MR_HIP(hipMallocAsync((void **)&d_dl, sizeof(float) * n_proj, s));
        hipError_t e = hipMemcpyAsync(d_dl, dl, sizeof(float) * n_proj, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipFreeAsync(d_dl, s);
            return mr::hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
        }
    }
    const size_t groups = (n_pix & 3) == 0 ? n_pix / 4 : n_pix;
    const unsigned bx = (unsigned)std::max<size_t>(1, std::min<size_t>((groups + 255) / 256, 64));
    hipError_t e;
    {
        mr::ProfScope prof("lognorm", s);
        hipLaunchKernelGGL(mr::lognorm_kernel, dim3(bx, (unsigned)n_proj), dim3(256), 0, s, d_proj, n_pix,
                           d_flat_log, d_dark, d_dl);
        e = hipGetLastError();
    }
    if (d_dl) (void)hipFreeAsync(d_dl, s);
    if (e != hipSuccess) return mr::hip_fail(e, "lognorm_kernel", __FILE__, __LINE__);
    // the host dose table must outlive the copy
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
}

}  // namespace

extern "C" int mr_lognorm(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                          const float *dose, float dark_dose, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(n_proj == 0 || (d_proj != nullptr && n_pix > 0), "NULL projection pointer or empty projection");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (n_proj == 0) return MR_OK;
    std::vector<float> dl;
    if (dose) {
        // ProcessCore(3D): doselist[i] = log(dose - fDarkDose < 1 ? 1 : dose - fDarkDose)
        dl.resize(n_proj);
        for (size_t j = 0; j < n_proj; ++j) {
            const float d = dose[j] - dark_dose;
            dl[j] = std::log(d < 1 ? 1.0f : d);
            MR_REQUIRE(dl[j] == dl[j], "The projection dose is a NaN");
        }
    }
    return run_lognorm(d_proj, n_proj, n_pix, d_flat_log, d_dark, dose ? dl.data() : nullptr, (hipStream_t)stream);
    MR_GUARD_END
}

namespace {

int run_lognorm_lut(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                    const float *term, hipStream_t s)
{
    const float *lut = nullptr;
    int rc = mr::log_lut_device(&lut);
    if (rc != MR_OK) return rc;
    float *d_t = nullptr;
    if (term) {
        MR_HIP(hipMallocAsync((void **)&d_t, sizeof(float) * n_proj, s));
        hipError_t e = hipMemcpyAsync(d_t, term, sizeof(float) * n_proj, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipFreeAsync(d_t, s);
            return mr::hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
        }
    }
    const size_t groups = (n_pix & 3) == 0 ? n_pix / 4 : n_pix;
    const unsigned bx = (unsigned)std::max<size_t>(1, std::min<size_t>((groups + 255) / 256, 64));
    hipError_t e;
    {
        mr::ProfScope prof("lognorm_lut", s);
        hipLaunchKernelGGL(mr::lognorm_lut_kernel, dim3(bx, (unsigned)n_proj), dim3(256), 0, s, d_proj, n_pix,
                           d_flat_log, d_dark, d_t, lut);
        e = hipGetLastError();
    }
    if (d_t) (void)hipFreeAsync(d_t, s);
    if (e != hipSuccess) return mr::hip_fail(e, "lognorm_lut_kernel", __FILE__, __LINE__);
    MR_HIP(hipStreamSynchronize(s));   // the host term table must outlive the copy
    return MR_OK;
}

}  // namespace

extern "C" int mr_lognorm_lut(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log,
                              const float *d_dark, const float *dose, float dark_dose, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(n_proj == 0 || (d_proj != nullptr && n_pix > 0), "NULL projection pointer or empty projection");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (n_proj == 0) return MR_OK;
    // ProcessCore(3D) with bUseLUT (NormPlugins.cpp:418-428, 455-475):
    // doselist[j] = log(max(dose[j] - fDarkDose, 1)) (0 without
    // usenormregion); without open beams the term is doselist[j] - doselist[0]
    std::vector<float> t(n_proj, 0.0f);
    if (dose) {
        for (size_t j = 0; j < n_proj; ++j) {
            const float d = dose[j] - dark_dose;
            t[j] = std::log(d < 1 ? 1.0f : d);
            MR_REQUIRE(t[j] == t[j], "The projection dose is a NaN");
        }
    }
    if (!d_flat_log) {
        const float d0 = t[0];
        for (size_t j = 0; j < n_proj; ++j) t[j] = t[j] - d0;
    }
    return run_lognorm_lut(d_proj, n_proj, n_pix, d_flat_log, d_dark, t.data(), (hipStream_t)stream);
    MR_GUARD_END
}

extern "C" int mr_lognorm_lut_term(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log,
                                   const float *d_dark, const float *term, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(n_proj == 0 || (d_proj != nullptr && n_pix > 0), "NULL projection pointer or empty projection");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (n_proj == 0) return MR_OK;
    return run_lognorm_lut(d_proj, n_proj, n_pix, d_flat_log, d_dark, term, (hipStream_t)stream);
    MR_GUARD_END
}

extern "C" int mr_lognorm_logdose(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log,
                                  const float *d_dark, const float *dose_log, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(n_proj == 0 || (d_proj != nullptr && n_pix > 0), "NULL projection pointer or empty projection");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (n_proj == 0) return MR_OK;
    return run_lognorm(d_proj, n_proj, n_pix, d_flat_log, d_dark, dose_log, (hipStream_t)stream);
    MR_GUARD_END
}
