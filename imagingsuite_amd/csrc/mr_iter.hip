// Iterative reconstruction on the matched Joseph pair (A, A^T): SIRT,
// Landweber and CGLS as sketched in the reference's
// frameworks/tomography/Backprojectors/IterativeBackProj/src/IterativeMeth.txt
// (landweberOTF :5-17, SIRT_OTF :21-52, cgLS_OTF :56-88); the reference's own
// SIRTbp::Process (.../src/sirtbp.cpp:32-81) is an empty stub.
//
//   SIRT:      R = 1 / A 1 (0 where A 1 = 0), C = 1 / A^T 1 (likewise)
//              x <- x + C . A^T (R . (b - A x))
//   Landweber: x <- x + tau A^T (b - A x), then x <- max(x, 0) when clamped
//   CGLS:      r = b - A x, d = A^T r, |s|^2 = d.d; per iteration
//              q = A d, alpha = |s|^2 / q.q, x += alpha d, r -= alpha q,
//              s = A^T r, beta = s.s / |s|^2_old, d = s + beta d
//
// Every iteration is enqueued on one stream with no host synchronisation:
// the CGLS scalars live on the device and the dot products are two-pass
// reductions with a fixed order (deterministic results).  The residual
// history |b - A x|^2 (before each update) comes from the same passes.
#include "mr_common.h"
#include "mr_filter.h"
#include "mr_joseph.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mr {

constexpr int kRedBlocks = 512;   // partial sums per reduction
constexpr int kRedThreads = 256;

enum { SC_NORMR2 = 0, SC_ALPHA = 1, SC_BETA = 2, SC_COUNT = 4 };
enum { FIN_STORE = 0, FIN_ALPHA = 1, FIN_BETA = 2 };

__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kRedThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kRedThreads / 64; ++i) t += red[i];
    return t;
}

// part[b] = sum over the block's grid-stride elements of a[i] * b[i]
__global__ __launch_bounds__(kRedThreads) void dot_partial_kernel(const float *a, const float *b, size_t n,
                                                                  double *part)
{
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads)
        s += (double)a[i] * (double)b[i];
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sums the kRedBlocks partials in a fixed order; op decides where it goes:
// FIN_STORE dst[0] = sum; FIN_ALPHA sc[ALPHA] = sc[NORMR2] / sum;
// FIN_BETA sc[BETA] = sum / sc[NORMR2], sc[NORMR2] = sum.
__global__ __launch_bounds__(kRedThreads) void reduce_final_kernel(const double *part, double *dst, double *sc,
                                                                   int op)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) s += part[i];
    s = block_sum(s);
    if (threadIdx.x != 0) return;
    if (op == FIN_STORE) {
        dst[0] = s;
    } else if (op == FIN_ALPHA) {
        sc[SC_ALPHA] = s > 0.0 ? sc[SC_NORMR2] / s : 0.0;
    } else {
        sc[SC_BETA] = sc[SC_NORMR2] > 0.0 ? s / sc[SC_NORMR2] : 0.0;
        sc[SC_NORMR2] = s;
    }
}

// r = b - r (r holds A x on entry), partial sums of r^2 when part != NULL,
// then r *= R when R != NULL.
__global__ __launch_bounds__(kRedThreads) void residual_kernel(float *r, const float *b, const float *R, size_t n,
                                                               double *part)
{
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        const float v = b[i] - r[i];
        s += (double)v * (double)v;
        r[i] = R ? v * R[i] : v;
    }
    if (part) {
        s = block_sum(s);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// x += c d with c = C[i] (SIRT), tau (Landweber) or sign * alpha (CGLS);
// clamped at zero when nonneg.
__global__ __launch_bounds__(256) void update_kernel(float *x, const float *d, const float *C, float tau,
                                                     const double *sc, float sign, int nonneg, size_t n)
{
    const float ca = sc ? sign * (float)sc[SC_ALPHA] : tau;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float c = C ? C[i] : ca;
        float v = x[i] + c * d[i];
        if (nonneg) v = fmaxf(v, 0.0f);
        x[i] = v;
    }
}

// d = s + beta d
__global__ __launch_bounds__(256) void xpby_kernel(float *d, const float *s, const double *sc, size_t n)
{
    const float beta = (float)sc[SC_BETA];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = s[i] + beta * d[i];
}

__global__ __launch_bounds__(256) void fill_kernel(float *v, float value, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = value;
}

// v = v > 0 ? 1 / v : 0
__global__ __launch_bounds__(256) void recip_kernel(float *v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = v[i] > 0.0f ? 1.0f / v[i] : 0.0f;
}

unsigned ew_blocks(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

}  // namespace mr

struct mr_iter_plan {
    int device = 0;
    mr_joseph_geom g{};
    size_t n_ang = 0, nvol = 0, nsino = 0;
    int method = 0;
    void *d_tabs = nullptr;                 // packed angle tables
    mr::JosephDevTables dt;
    float *R = nullptr, *C = nullptr;      // SIRT weights
    float *r = nullptr, *q = nullptr;      // sinogram work (q: CGLS)
    float *v1 = nullptr, *v2 = nullptr;    // volume work (v2: CGLS)
    double *part = nullptr, *sc = nullptr;
};

namespace {

void free_plan(mr_iter_plan *p)
{
    if (!p) return;
    (void)hipDeviceSynchronize();
    for (void *ptr : {p->d_tabs, (void *)p->R, (void *)p->C, (void *)p->r, (void *)p->q, (void *)p->v1,
                      (void *)p->v2, (void *)p->part, (void *)p->sc})
        if (ptr) (void)hipFree(ptr);
    delete p;
}

int dot(mr_iter_plan *p, const float *a, const float *b, size_t n, double *dst, int op, hipStream_t s)
{
    hipLaunchKernelGGL(mr::dot_partial_kernel, dim3(mr::kRedBlocks), dim3(mr::kRedThreads), 0, s, a, b, n, p->part);
    hipLaunchKernelGGL(mr::reduce_final_kernel, dim3(1), dim3(mr::kRedThreads), 0, s, p->part, dst, p->sc, op);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// r = b - A x (into p->r), optional |r|^2 into hist, r *= R (SIRT)
int residual(mr_iter_plan *p, const float *b, const float *x, const float *R, double *hist, hipStream_t s)
{
    int rc = mr::joseph_launch(&p->g, x, p->r, p->dt, p->n_ang, true, s);
    if (rc != MR_OK) return rc;
    hipLaunchKernelGGL(mr::residual_kernel, dim3(mr::kRedBlocks), dim3(mr::kRedThreads), 0, s, p->r, b, R, p->nsino,
                       hist ? p->part : nullptr);
    if (hist)
        hipLaunchKernelGGL(mr::reduce_final_kernel, dim3(1), dim3(mr::kRedThreads), 0, s, p->part, hist, p->sc,
                           (int)mr::FIN_STORE);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace

extern "C" int mr_iter_plan_create(mr_iter_plan **plan, const mr_joseph_geom *geom, const float *angles, size_t n_ang,
                                   int method)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(plan != nullptr && angles != nullptr, "NULL argument");
    *plan = nullptr;
    int rc = mr::joseph_check(geom, n_ang);
    if (rc != MR_OK) return rc;
    MR_REQUIRE(method == MR_ITER_SIRT || method == MR_ITER_LANDWEBER || method == MR_ITER_CGLS,
               "method must be MR_ITER_SIRT, MR_ITER_LANDWEBER or MR_ITER_CGLS");
    auto *p = new mr_iter_plan();
    p->g = *geom;
    p->n_ang = n_ang;
    p->method = method;
    p->nvol = geom->nx * geom->nx * geom->nz;
    p->nsino = n_ang * geom->n_det * geom->nz;
    if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
    const mr::JosephPacked pk = mr::joseph_pack(*geom, angles, n_ang);
    hipError_t e = hipMalloc(&p->d_tabs, pk.bytes());
    if (e == hipSuccess) {
        rc = mr::joseph_upload(pk, p->d_tabs, nullptr, p->dt);
        e = rc == MR_OK ? hipStreamSynchronize(nullptr) : hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipMalloc(&p->r, sizeof(float) * p->nsino);
    if (e == hipSuccess) e = hipMalloc(&p->v1, sizeof(float) * p->nvol);
    if (e == hipSuccess) e = hipMalloc(&p->part, sizeof(double) * mr::kRedBlocks);
    if (e == hipSuccess) e = hipMalloc(&p->sc, sizeof(double) * mr::SC_COUNT);
    if (e == hipSuccess && method == MR_ITER_SIRT) {
        e = hipMalloc(&p->R, sizeof(float) * p->nsino);
        if (e == hipSuccess) e = hipMalloc(&p->C, sizeof(float) * p->nvol);
    }
    if (e == hipSuccess && method == MR_ITER_CGLS) {
        e = hipMalloc(&p->q, sizeof(float) * p->nsino);
        if (e == hipSuccess) e = hipMalloc(&p->v2, sizeof(float) * p->nvol);
    }
    if (e != hipSuccess) {
        free_plan(p);
        return mr::hip_fail(e, "mr_iter_plan_create", __FILE__, __LINE__);
    }
    if (method == MR_ITER_SIRT) {
        // R = 1 / A 1, C = 1 / A^T 1
        hipStream_t s = nullptr;
        hipLaunchKernelGGL(mr::fill_kernel, dim3(mr::ew_blocks(p->nvol)), dim3(256), 0, s, p->v1, 1.0f, p->nvol);
        rc = mr::joseph_launch(&p->g, p->v1, p->R, p->dt, n_ang, true, s);
        if (rc == MR_OK) {
            hipLaunchKernelGGL(mr::fill_kernel, dim3(mr::ew_blocks(p->nsino)), dim3(256), 0, s, p->r, 1.0f, p->nsino);
            rc = mr::joseph_launch(&p->g, p->r, p->C, p->dt, n_ang, false, s);
        }
        if (rc == MR_OK) {
            hipLaunchKernelGGL(mr::recip_kernel, dim3(mr::ew_blocks(p->nsino)), dim3(256), 0, s, p->R, p->nsino);
            hipLaunchKernelGGL(mr::recip_kernel, dim3(mr::ew_blocks(p->nvol)), dim3(256), 0, s, p->C, p->nvol);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = mr::hip_fail(e, "SIRT weights", __FILE__, __LINE__);
        }
        if (rc != MR_OK) {
            free_plan(p);
            return rc;
        }
    }
    *plan = p;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_iter_plan_destroy(mr_iter_plan *plan)
{
    MR_GUARD_BEGIN
    free_plan(plan);
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_iter_run(mr_iter_plan *p, const float *d_sino, float *d_volume, int iterations, float tau,
                           int nonneg, double *residuals, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(p != nullptr && d_sino != nullptr && d_volume != nullptr, "NULL argument");
    MR_REQUIRE(iterations >= 0, "iterations must be >= 0");
    MR_REQUIRE(p->method != MR_ITER_LANDWEBER || (tau > 0 && std::isfinite(tau)),
               "Landweber needs a finite step tau > 0");
    MR_REQUIRE(nonneg >= -1 && nonneg <= 1, "nonneg must be -1 (method default), 0 or 1");
    if (nonneg < 0) nonneg = p->method == MR_ITER_LANDWEBER ? 1 : 0;   // landweberOTF clamps, SIRT_OTF does not
    MR_REQUIRE(!(p->method == MR_ITER_CGLS && nonneg), "CGLS has no non-negativity clamp");
    if (iterations == 0) return MR_OK;
    hipStream_t s = (hipStream_t)stream;
    double *d_hist = nullptr;
    if (residuals) MR_HIP(hipMallocAsync((void **)&d_hist, sizeof(double) * (size_t)iterations, s));
    const unsigned bv = mr::ew_blocks(p->nvol);
    int rc = MR_OK;
    if (p->method == MR_ITER_SIRT || p->method == MR_ITER_LANDWEBER) {
        const bool sirt = p->method == MR_ITER_SIRT;
        for (int k = 0; k < iterations && rc == MR_OK; ++k) {
            rc = residual(p, d_sino, d_volume, sirt ? p->R : nullptr, d_hist ? d_hist + k : nullptr, s);
            if (rc == MR_OK) rc = mr::joseph_launch(&p->g, p->r, p->v1, p->dt, p->n_ang, false, s);
            if (rc == MR_OK) {
                hipLaunchKernelGGL(mr::update_kernel, dim3(bv), dim3(256), 0, s, d_volume, p->v1,
                                   sirt ? p->C : nullptr, tau, nullptr, 1.0f, nonneg, p->nvol);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) rc = mr::hip_fail(e, "update_kernel", __FILE__, __LINE__);
            }
        }
    } else {
        // CGLS: d (v2) = A^T r, |s|^2 = d.d
        rc = residual(p, d_sino, d_volume, nullptr, nullptr, s);
        if (rc == MR_OK) rc = mr::joseph_launch(&p->g, p->r, p->v2, p->dt, p->n_ang, false, s);
        if (rc == MR_OK) rc = dot(p, p->v2, p->v2, p->nvol, p->sc + mr::SC_NORMR2, mr::FIN_STORE, s);
        const unsigned bs = mr::ew_blocks(p->nsino);
        for (int k = 0; k < iterations && rc == MR_OK; ++k) {
            if (d_hist) rc = dot(p, p->r, p->r, p->nsino, d_hist + k, mr::FIN_STORE, s);
            if (rc == MR_OK) rc = mr::joseph_launch(&p->g, p->v2, p->q, p->dt, p->n_ang, true, s);
            if (rc == MR_OK) rc = dot(p, p->q, p->q, p->nsino, nullptr, mr::FIN_ALPHA, s);
            if (rc != MR_OK) break;
            hipLaunchKernelGGL(mr::update_kernel, dim3(bv), dim3(256), 0, s, d_volume, p->v2, nullptr, 0.0f, p->sc,
                               1.0f, 0, p->nvol);
            hipLaunchKernelGGL(mr::update_kernel, dim3(bs), dim3(256), 0, s, p->r, p->q, nullptr, 0.0f, p->sc,
                               -1.0f, 0, p->nsino);
            rc = mr::joseph_launch(&p->g, p->r, p->v1, p->dt, p->n_ang, false, s);
            if (rc == MR_OK) rc = dot(p, p->v1, p->v1, p->nvol, nullptr, mr::FIN_BETA, s);
            if (rc != MR_OK) break;
            hipLaunchKernelGGL(mr::xpby_kernel, dim3(bv), dim3(256), 0, s, p->v2, p->v1, p->sc, p->nvol);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = mr::hip_fail(e, "CGLS update", __FILE__, __LINE__);
        }
    }
    if (d_hist) {
        if (rc == MR_OK) {
            hipError_t e = hipMemcpyAsync(residuals, d_hist, sizeof(double) * (size_t)iterations,
                                          hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = mr::hip_fail(e, "residual history", __FILE__, __LINE__);
        }
        (void)hipFreeAsync(d_hist, s);
    }
    return rc;
    MR_GUARD_END
}
