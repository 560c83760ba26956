// Streaming parallel-beam FBP of a host-resident projection stack.
//
// The reference engine reconstructs a stack block by block
// (ReconEngine::Run3DFull / Process3D,
// frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:838-1185
// and 1314-1714): for each SliceBlock of detector rows it copies the rows of
// every projection out of the stack, runs the pre-processing chain
// (ProjectionFilterSingle) and the back-projector (MultiProjBPparallel), and
// copies the block's slices into the output volume (TransferMatrix,
// :531-574).  Here the same blocks go through one GPU pipeline:
//
//   host rows --H2D--> d_in[i] --filter, weight, BP--> d_vol --transpose--> d_out[i] --D2H--> host slices
//
// with two slab buffers per direction, one stream for host-to-device copies,
// one for the kernels and one for device-to-host copies, so the copies of
// block b+1 and b-1 run under the kernels of block b.  The projections stay on
// the GPU between filter and back-projection.  Host buffers allocated with
// mr_host_alloc (pinned) are copied by DMA directly (2-D copies gather the
// block's rows); other host memory goes through pinned staging buffers filled
// by host threads.  The output is slice-major [rows][width][width], the order
// of BackProjectorModuleBase::GetSlice.
#include "mr_bp.h"
#include "mr_common.h"
#include "mr_filter.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace mr {

// out[z][y][x] = vol[y][x][z] for z < nz (one y row, 64 x, 32 z per block)
__global__ __launch_bounds__(256) void zxy_to_slices_kernel(const float *vol, float *out, int w, int nz)
{
    __shared__ float t[64][33];
    const int x0 = blockIdx.x * 64, y = blockIdx.y, z0 = blockIdx.z * 32;
    const float *src = vol + (size_t)y * w * nz;
    // read: for each x, 32 consecutive z
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int xl = e >> 5, zl = e & 31;
        const int x = x0 + xl, z = z0 + zl;
        t[xl][zl] = (x < w && z < nz) ? src[(size_t)x * nz + z] : 0.0f;
    }
    __syncthreads();
    // write: for each z, 64 consecutive x
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int zl = e >> 6, xl = e & 63;
        const int x = x0 + xl, z = z0 + zl;
        if (x < w && z < nz) out[((size_t)z * w + y) * w + x] = t[xl][zl];
    }
}

// Pinned allocations made through mr_host_alloc (start -> bytes).
std::mutex g_pin_mutex;
std::map<const char *, size_t> g_pinned;

bool is_pinned(const void *p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_pin_mutex);
    const char *c = static_cast<const char *>(p);
    auto it = g_pinned.upper_bound(c);
    if (it == g_pinned.begin()) return false;
    --it;
    return c >= it->first && c + bytes <= it->first + it->second;
}

// Runs f(t, n_threads) on n_threads host threads (the calling one included).
template <class F>
void parallel_for_threads(int n_threads, F f)
{
    std::vector<std::thread> th;
    for (int t = 1; t < n_threads; ++t) th.emplace_back(f, t, n_threads);
    f(0, n_threads);
    for (auto &x : th) x.join();
}

}  // namespace mr

struct mr_fbp_stream {
    int device = 0;
    size_t width = 0, rows = 0, n_proj = 0, sb = 0;
    mr_filter_plan *fplan = nullptr;
    mr_bp_plan *bp_full = nullptr, *bp_rem = nullptr;   // slab heights sb and rows % sb
    float *d_in[2] = {nullptr, nullptr};
    float *d_vol = nullptr;
    float *d_out[2] = {nullptr, nullptr};
    float *h_in[2] = {nullptr, nullptr};                 // pinned staging (pageable callers)
    float *h_out[2] = {nullptr, nullptr};
    hipStream_t h2d = nullptr, cs = nullptr, d2h = nullptr;
    hipEvent_t ev_in[2] = {nullptr, nullptr}, ev_comp[2] = {nullptr, nullptr}, ev_out[2] = {nullptr, nullptr};
    int n_threads = 8;
    size_t tilt_row0 = 0;   // detector row of the first slab (tilt correction)
};

namespace {

void destroy_stream(mr_fbp_stream *s)
{
    if (!s) return;
    (void)hipDeviceSynchronize();
    if (s->fplan) mr_filter_plan_destroy(s->fplan);
    if (s->bp_full) mr_bp_plan_destroy(s->bp_full);
    if (s->bp_rem) mr_bp_plan_destroy(s->bp_rem);
    for (int i = 0; i < 2; ++i) {
        if (s->d_in[i]) (void)hipFree(s->d_in[i]);
        if (s->d_out[i]) (void)hipFree(s->d_out[i]);
        if (s->h_in[i]) (void)hipHostFree(s->h_in[i]);
        if (s->h_out[i]) (void)hipHostFree(s->h_out[i]);
        if (s->ev_in[i]) (void)hipEventDestroy(s->ev_in[i]);
        if (s->ev_comp[i]) (void)hipEventDestroy(s->ev_comp[i]);
        if (s->ev_out[i]) (void)hipEventDestroy(s->ev_out[i]);
    }
    if (s->d_vol) (void)hipFree(s->d_vol);
    for (hipStream_t st : {s->h2d, s->cs, s->d2h})
        if (st) (void)hipStreamDestroy(st);
    delete s;
}

}  // namespace

extern "C" int mr_host_alloc(void **ptr, size_t bytes)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(ptr != nullptr && bytes > 0, "NULL pointer or zero size");
    *ptr = nullptr;
    MR_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    std::lock_guard<std::mutex> lk(mr::g_pin_mutex);
    mr::g_pinned[static_cast<const char *>(*ptr)] = bytes;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_host_free(void *ptr)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    if (!ptr) return MR_OK;
    {
        std::lock_guard<std::mutex> lk(mr::g_pin_mutex);
        auto it = mr::g_pinned.find(static_cast<const char *>(ptr));
        MR_REQUIRE(it != mr::g_pinned.end(), "pointer was not allocated by mr_host_alloc");
        mr::g_pinned.erase(it);
    }
    MR_HIP(hipHostFree(ptr));
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fbp_stream_create(mr_fbp_stream **stream, const mr_bp_params *bp, size_t rows, size_t n_proj,
                                    size_t slice_block, int filter_type, float cutoff, float order, int use_bias,
                                    float bias_weight, size_t padding_doubler)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(stream != nullptr && bp != nullptr, "NULL argument");
    *stream = nullptr;
    MR_REQUIRE(rows > 0 && n_proj > 0 && slice_block > 0, "rows, projections and slice block must be > 0");
    MR_REQUIRE(bp->width >= 2, "width must be >= 2");
    auto *s = new mr_fbp_stream();
    s->width = bp->width;
    s->rows = rows;
    s->n_proj = n_proj;
    s->sb = std::min(slice_block, rows);
    const unsigned hc = std::thread::hardware_concurrency();
    s->n_threads = (int)std::max(1u, std::min(hc == 0 ? 8u : hc, 16u));
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    int rc = mr_filter_plan_create(&s->fplan, bp->width, filter_type, cutoff, order, use_bias, bias_weight,
                                   padding_doubler);
    mr_bp_params q = *bp;
    s->tilt_row0 = bp->tilt_row0;
    if (rc == MR_OK) {
        q.slices = s->sb;
        rc = mr_bp_plan_create(&s->bp_full, &q);
    }
    if (rc == MR_OK && rows % s->sb) {
        q.slices = rows % s->sb;
        rc = mr_bp_plan_create(&s->bp_rem, &q);
    }
    if (rc != MR_OK) {
        destroy_stream(s);
        return rc;
    }
    const size_t w = s->width, sb = s->sb;
    const size_t in_b = sizeof(float) * n_proj * sb * w, vol_b = sizeof(float) * w * w * sb;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipMalloc(&s->d_in[i], in_b);
        if (e == hipSuccess) e = hipMalloc(&s->d_out[i], vol_b);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_in[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_comp[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_out[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipMalloc(&s->d_vol, vol_b);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->h2d, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->d2h, hipStreamNonBlocking);
    if (e != hipSuccess) {
        destroy_stream(s);
        return mr::hip_fail(e, "mr_fbp_stream_create", __FILE__, __LINE__);
    }
    *stream = s;
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_fbp_stream_destroy(mr_fbp_stream *stream)
{
    MR_GUARD_BEGIN
    destroy_stream(stream);
    return MR_OK;
    MR_GUARD_END
}

namespace {

int staging(mr_fbp_stream *s, bool in)
{
    const size_t w = s->width, sb = s->sb;
    float **h = in ? s->h_in : s->h_out;
    const size_t bytes = in ? sizeof(float) * s->n_proj * sb * w : sizeof(float) * w * w * sb;
    for (int i = 0; i < 2; ++i)
        if (!h[i]) MR_HIP(hipHostMalloc((void **)&h[i], bytes, hipHostMallocDefault));
    return MR_OK;
}

int run_stream(mr_fbp_stream *s, const float *h_proj, const float *angles, const float *weights, float *h_vol,
               bool serial, mr_slab_fn on_slab = nullptr, void *user = nullptr)
{
    const size_t w = s->width, rows = s->rows, np = s->n_proj, sb = s->sb;
    const bool pin_in = mr::is_pinned(h_proj, sizeof(float) * np * rows * w);
    const bool pin_out = mr::is_pinned(h_vol, sizeof(float) * rows * w * w);
    int rc = MR_OK;
    if (!pin_in && (rc = staging(s, true)) != MR_OK) return rc;
    if (!pin_out && (rc = staging(s, false)) != MR_OK) return rc;
    if ((rc = mr::bp_load_tables(s->bp_full, np, angles, weights, s->cs)) != MR_OK) return rc;
    if (s->bp_rem && (rc = mr::bp_load_tables(s->bp_rem, np, angles, weights, s->cs)) != MR_OK) return rc;
    const size_t nb = (rows + sb - 1) / sb;
    const int T = s->n_threads;
    auto scatter = [&](size_t b) -> int {
        // slab b of the staging buffer into the caller's volume
        const int i = (int)(b & 1);
        const size_t z0 = b * sb, nr = std::min(sb, rows - z0);
        MR_HIP(hipEventSynchronize(s->ev_out[i]));
        const size_t n = nr * w * w;
        mr::parallel_for_threads(T, [&](int t, int nt) {
            const size_t a = n * (size_t)t / (size_t)nt, e = n * (size_t)(t + 1) / (size_t)nt;
            std::memcpy(h_vol + z0 * w * w + a, s->h_out[i] + a, sizeof(float) * (e - a));
        });
        return MR_OK;
    };
    // slab b is in the caller's volume: hand it over (the GPU works on the
    // later slabs meanwhile); a nonzero return stops the run
    auto deliver = [&](size_t b) -> int {
        if (!on_slab) return MR_OK;
        const size_t z0 = b * sb, nr = std::min(sb, rows - z0);
        if (pin_out) MR_HIP(hipEventSynchronize(s->ev_out[b & 1]));
        if (on_slab(user, z0, nr, h_vol + z0 * w * w) != 0) {
            mr::set_error("the slab callback stopped the reconstruction");
            return MR_ERR_ABORTED;
        }
        return MR_OK;
    };
    for (size_t b = 0; b < nb; ++b) {
        const int i = (int)(b & 1);
        const size_t z0 = b * sb, nr = std::min(sb, rows - z0);
        mr_bp_plan *plan = nr == sb ? s->bp_full : s->bp_rem;
        // host -> device: the block's rows of every projection
        const float *src;
        size_t spitch;
        if (pin_in) {
            src = h_proj + z0 * w;
            spitch = sizeof(float) * rows * w;
        } else {
            if (b >= 2) MR_HIP(hipEventSynchronize(s->ev_in[i]));   // staging buffer i is free
            float *dst = s->h_in[i];
            mr::parallel_for_threads(T, [&](int t, int nt) {
                for (size_t p = (size_t)t; p < np; p += (size_t)nt)
                    std::memcpy(dst + p * nr * w, h_proj + (p * rows + z0) * w, sizeof(float) * nr * w);
            });
            src = dst;
            spitch = sizeof(float) * nr * w;
        }
        if (b >= 2) MR_HIP(hipStreamWaitEvent(s->h2d, s->ev_comp[i], 0));   // d_in[i] consumed
        MR_HIP(hipMemcpy2DAsync(s->d_in[i], sizeof(float) * nr * w, src, spitch, sizeof(float) * nr * w, np,
                                hipMemcpyHostToDevice, s->h2d));
        MR_HIP(hipEventRecord(s->ev_in[i], s->h2d));
        // kernels
        MR_HIP(hipStreamWaitEvent(s->cs, s->ev_in[i], 0));
        if ((rc = mr_filter_rows(s->fplan, s->d_in[i], np * nr, s->cs)) != MR_OK) return rc;
        // each slab is a block of its own: the tilt terms use its first row
        mr::bp_set_tilt_row0(plan, s->tilt_row0 + z0);
        if ((rc = mr::bp_enqueue(plan, s->d_in[i], s->d_vol, 0, s->cs)) != MR_OK) return rc;
        if (b >= 2) MR_HIP(hipStreamWaitEvent(s->cs, s->ev_out[i], 0));    // d_out[i] copied out
        {
            dim3 g((unsigned)((w + 63) / 64), (unsigned)w, (unsigned)((nr + 31) / 32));
            mr::ProfScope prof("zxy_to_slices", s->cs);
            hipLaunchKernelGGL(mr::zxy_to_slices_kernel, g, dim3(256), 0, s->cs, s->d_vol, s->d_out[i], (int)w,
                               (int)nr);
            MR_HIP(hipGetLastError());
        }
        MR_HIP(hipEventRecord(s->ev_comp[i], s->cs));
        // device -> host: the block's slices
        MR_HIP(hipStreamWaitEvent(s->d2h, s->ev_comp[i], 0));
        float *dst = pin_out ? h_vol + z0 * w * w : s->h_out[i];
        MR_HIP(hipMemcpyAsync(dst, s->d_out[i], sizeof(float) * nr * w * w, hipMemcpyDeviceToHost, s->d2h));
        MR_HIP(hipEventRecord(s->ev_out[i], s->d2h));
        if (!serial && !pin_out && b >= 1 && (rc = scatter(b - 1)) != MR_OK) return rc;
        if (!serial && b >= 1 && (rc = deliver(b - 1)) != MR_OK) return rc;
        if (serial) {
            MR_HIP(hipStreamSynchronize(s->d2h));
            if (!pin_out && (rc = scatter(b)) != MR_OK) return rc;
            if ((rc = deliver(b)) != MR_OK) return rc;
        }
    }
    if (!serial && !pin_out && (rc = scatter(nb - 1)) != MR_OK) return rc;
    if (!serial && (rc = deliver(nb - 1)) != MR_OK) return rc;
    MR_HIP(hipStreamSynchronize(s->d2h));
    MR_HIP(hipStreamSynchronize(s->cs));
    MR_HIP(hipStreamSynchronize(s->h2d));
    return MR_OK;
}

}  // namespace

extern "C" int mr_fbp_stream_run_slabs(mr_fbp_stream *stream, const float *h_proj, const float *angles,
                                       const float *weights, float *h_volume, mr_slab_fn on_slab, void *user)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(stream != nullptr && h_proj != nullptr && angles != nullptr && weights != nullptr &&
                   h_volume != nullptr && on_slab != nullptr,
               "NULL argument");
    int rc = run_stream(stream, h_proj, angles, weights, h_volume, false, on_slab, user);
    if (rc != MR_OK) {
        (void)hipStreamSynchronize(stream->h2d);
        (void)hipStreamSynchronize(stream->cs);
        (void)hipStreamSynchronize(stream->d2h);
    }
    return rc;
    MR_GUARD_END
}

extern "C" int mr_fbp_stream_run(mr_fbp_stream *stream, const float *h_proj, const float *angles,
                                 const float *weights, float *h_volume, int serial)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(stream != nullptr && h_proj != nullptr && angles != nullptr && weights != nullptr &&
                   h_volume != nullptr,
               "NULL argument");
    int rc = run_stream(stream, h_proj, angles, weights, h_volume, serial != 0);
    if (rc != MR_OK) {
        // leave no copy or kernel of this call behind
        (void)hipStreamSynchronize(stream->h2d);
        (void)hipStreamSynchronize(stream->cs);
        (void)hipStreamSynchronize(stream->d2h);
    }
    return rc;
    MR_GUARD_END
}
