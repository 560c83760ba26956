// Back-projection that allocates nothing (mr_bp_parallel_noalloc).
//
// Callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation) get the parallel-beam
// back-projection of mr_bp_parallel without a plan: the circle mask, the
// projection tables and the weighted, transposed sinogram live in one
// file-scope device array, `scratch`, of MR_NOALLOC_SCRATCH_BYTES.
//
//   scratch layout (bytes)
//     [0, kTableBytes)             mask int2[width] | tab float4[n_used] |
//                                  idx int[n_used] | w float[n_proj], each
//                                  part 256-byte aligned
//     [kTableBytes, kScratchBytes) sinogram chunk [kc][width][vpad]
//
// The used projections are back-projected in chunks of kc = as many staged
// projections as fit; each chunk is weighted/transposed into the scratch and
// added to the volume by the same kernels as mr_bp_parallel.  Every voxel
// still accumulates the projections one after the other in the reference's
// order (the kernels load the voxel, add the chunk's projections in order
// and store it), so the result equals mr_bp_parallel's bit for bit.
//
// The host tables reach the scratch through kernel arguments (scratch_put_kernel),
// so no host buffer has to outlive the call and nothing is allocated.
//
// Each device holds its own copy of `scratch` at its own address, found once
// per device with hipGetSymbolAddress; host code only ever uses that address,
// never the array's name.  All calls on one device share the array: calls
// that overlap on one device (different streams, or threads) race.
#include "mr_bp.h"
#include "mr_common.h"
#include "mr_host.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr size_t kScratchBytes = MR_NOALLOC_SCRATCH_BYTES;
constexpr size_t kTableBytes = size_t(4) << 20;   // mask + tables (at most about 2.3 MiB)
constexpr int kMaxDevices = 64;
constexpr unsigned kPutWords = 480;               // words per scratch_put_kernel launch (1.9 KB of arguments)

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (current), looked up once.
int scratch_address(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        MR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MR_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a NULL scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return MR_OK;
}

struct PutArgs {
    unsigned off;              // first word written
    unsigned n;                // words written (<= kPutWords)
    unsigned w[kPutWords];
};

// base[off + i] = w[i]: moves a piece of a host table into the scratch by
// value, through the launch's arguments.
__global__ __launch_bounds__(256) void scratch_put_kernel(unsigned *__restrict__ base, PutArgs a)
{
    for (unsigned i = threadIdx.x; i < a.n; i += 256) base[a.off + i] = a.w[i];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" int mr_bp_parallel_noalloc(const mr_bp_params *params, const float *d_proj, size_t n_proj,
                                      const float *angles, const float *weights, float *d_volume, int accumulate,
                                      void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(params != nullptr, "params is NULL");
    MR_REQUIRE(d_volume != nullptr, "volume pointer is NULL");
    MR_REQUIRE(n_proj == 0 || (d_proj && angles && weights), "NULL projection/angle/weight pointer");
    const mr_bp_params &q = *params;
    MR_REQUIRE(q.width >= 2 && q.width <= 65536, "width must be in [2, 65536]");
    MR_REQUIRE(q.slices >= 1 && q.slices <= 65536, "slices must be in [1, 65536]");
    MR_REQUIRE(q.interp == MR_INTERP_LINEAR || q.interp == MR_INTERP_NEAREST, "unknown interpolation");
    MR_REQUIRE(q.direction == 0 || q.direction == 1, "direction must be 0 (CW) or 1 (CCW)");
    MR_REQUIRE(q.buffer_size >= 0, "buffer_size must be >= 0");
    if (q.use_roi) MR_REQUIRE(q.roi[0] <= q.roi[2] && q.roi[1] <= q.roi[3], "matrix ROI x1<x0 or y1<y0");
    MR_REQUIRE(n_proj <= 65535, "at most 65535 projections per call");
    if (q.correct_tilt) {
        mr::set_error("tilt correction runs through mr_bp_parallel (plan), not the no-allocation entry point");
        return MR_ERR_UNSUPPORTED;
    }
    const size_t vpad = (q.slices + 3) & ~size_t(3);
    const size_t proj_bytes = q.width * vpad * sizeof(float);
    const size_t sino_bytes = kScratchBytes - kTableBytes;
    if (proj_bytes > sino_bytes) {
        mr::set_error("mr_bp_parallel_noalloc: one staged projection (width x slices rounded up to 4 x 4 bytes = " +
                      std::to_string(proj_bytes) + " bytes) exceeds the " + std::to_string(sino_bytes) +
                      "-byte sinogram part of the fixed scratch; use mr_bp_parallel");
        return MR_ERR_INVALID;
    }
    const std::vector<size_t> used = mr::used_projections(n_proj, (size_t)q.buffer_size);
    const size_t nk = used.size();

    // table region: mask | tab | idx | w
    const size_t off_tab = align256(sizeof(int2) * q.width);
    const size_t off_idx = off_tab + align256(sizeof(float4) * nk);
    const size_t off_w = off_idx + align256(sizeof(int) * nk);
    const size_t table_end = off_w + sizeof(float) * n_proj;
    MR_REQUIRE(table_end <= kTableBytes, "tables exceed the fixed scratch");   // cannot happen with the limits above

    int dev = 0;
    MR_HIP(hipGetDevice(&dev));
    MR_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    unsigned char *base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc != MR_OK) return rc;
    hipStream_t s = (hipStream_t)stream;

    // host image of the table region, then into the scratch by kernel arguments
    std::vector<unsigned> img((table_end + 3) / 4, 0u);
    {
        unsigned char *h = reinterpret_cast<unsigned char *>(img.data());
        std::vector<size_t> first, second;
        mr::circle_mask(q.width, q.width, q.use_roi != 0, q.roi, false, first, second);
        for (size_t y = 0; y < q.width; ++y) {
            // reference loop x = first+1 .. second (1-based) == [first, second) 0-based
            const int2 m = make_int2((int)first[y], (int)std::max(first[y], second[y]));
            std::memcpy(h + sizeof(int2) * y, &m, sizeof(int2));
        }
        if (nk > 0) {
            std::vector<float> tab;
            mr::bp_tables(angles, used.data(), nk, q.direction, q.center, q.width / 2, tab);
            std::memcpy(h + off_tab, tab.data(), sizeof(float) * tab.size());
            for (size_t k = 0; k < nk; ++k) {
                const int v = (int)used[k];
                std::memcpy(h + off_idx + sizeof(int) * k, &v, sizeof(int));
            }
            std::memcpy(h + off_w, weights, sizeof(float) * n_proj);
        }
    }
    {
        PutArgs pa;
        for (size_t w0 = 0; w0 < img.size(); w0 += kPutWords) {
            pa.off = (unsigned)w0;
            pa.n = (unsigned)std::min<size_t>(kPutWords, img.size() - w0);
            std::memcpy(pa.w, img.data() + w0, sizeof(unsigned) * pa.n);
            hipLaunchKernelGGL(scratch_put_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<unsigned *>(base), pa);
            MR_HIP(hipGetLastError());
        }
    }

    const int2 *d_mask = reinterpret_cast<const int2 *>(base);
    const float4 *d_tab = reinterpret_cast<const float4 *>(base + off_tab);
    const int *d_idx = reinterpret_cast<const int *>(base + off_idx);
    const float *d_w = reinterpret_cast<const float *>(base + off_w);
    float *d_sino = reinterpret_cast<float *>(base + kTableBytes);
    const size_t kc_max = sino_bytes / proj_bytes;

    mr::BpArgs a{};
    a.vol = d_volume;
    a.mask = d_mask;
    a.sino = d_sino;
    a.nu = (int)q.width;
    a.nv = (int)q.slices;
    a.vpad = (int)vpad;
    a.nz4 = (int)((q.slices / 4) * 4);
    a.bstride = q.multiprojbp ? a.nz4 : (int)vpad;
    const int impl = (q.multiprojbp && (q.slices & 3)) ? 0 : 1;
    // nk == 0 still runs one launch: it writes the zeroed (or unchanged) volume
    size_t kb = 0;
    do {
        const size_t kc = std::min(kc_max, nk - kb);
        if (kc > 0) {
            rc = mr::launch_bp_prep(d_proj, d_sino, d_idx + kb, d_w, (int)q.width, (int)q.slices, (int)vpad,
                                    (int)kc, s);
            if (rc != MR_OK) return rc;
        }
        a.tab = d_tab + kb;
        a.nk = (int)kc;
        a.accumulate = (accumulate || kb > 0) ? 1 : 0;
        rc = mr::run_bp(a, q.interp, impl, q.fused ? 0 : 1, s);
        if (rc != MR_OK) return rc;
        kb += kc;
    } while (kb < nk);
    return MR_OK;
    MR_GUARD_END
}
