// Copies between pageable host memory and the device, and the module
// volume's transposition.
//
// MuhRec's module API hands the modules host images (kipl TImage, plain
// pageable memory) and reads the volume back from the host
// (BackProjectorModuleBase::GetSlice is not virtual,
// frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp:62-124).
// A plain hipMemcpy of pageable memory runs at a fraction of the link's
// rate; here the bytes go through two pinned staging buffers per device and
// direction: host threads fill (or drain) one while the DMA engine moves the
// other, and an upload and a download on two streams run at once (the link
// is full duplex).
// Host memory from mr_host_alloc is copied by DMA directly.
#include "mr_common.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace mr {

bool is_pinned(const void *p, size_t bytes);   // mr_stream.hip

namespace {

constexpr size_t kChunk = size_t(64) << 20;   // bytes per staging buffer
constexpr int kMaxDevices = 64;

struct Staging {
    std::mutex mu;
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};   // last DMA from / into buf[i]
    bool used[2] = {false, false};
};

Staging g_staging[2][kMaxDevices];   // [0]: host to device, [1]: device to host

int copy_threads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc == 0 ? 8u : hc, 8u));
}

// memcpy of n bytes on up to copy_threads() threads (the caller's included)
void parallel_memcpy(char *dst, const char *src, size_t n)
{
    const int T = n >= (size_t(4) << 20) ? copy_threads() : 1;
    if (T == 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + T - 1) / T + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) {
        const size_t a = per * t, e = std::min(n, a + per);
        if (a < e) th.emplace_back([=] { std::memcpy(dst + a, src + a, e - a); });
    }
    std::memcpy(dst, src, std::min(n, per));
    for (auto &x : th) x.join();
}

int staging(int dir, Staging **out)
{
    int dev = 0;
    MR_HIP(hipGetDevice(&dev));
    MR_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the staging table");
    Staging &s = g_staging[dir][dev];
    if (!s.buf[0]) {
        for (int i = 0; i < 2; ++i) {
            MR_HIP(hipHostMalloc((void **)&s.buf[i], kChunk, hipHostMallocDefault));
            MR_HIP(hipEventCreateWithFlags(&s.done[i], hipEventDisableTiming));
        }
    }
    *out = &s;
    return MR_OK;
}

}  // namespace

// out[z][y][x] = vol[y][x][z] for z < nz (one y row, 64 x, 32 z per block)
__global__ __launch_bounds__(256) void zxy_to_xyz_kernel(const float *__restrict__ vol, float *__restrict__ out,
                                                         int w, int nz)
{
    __shared__ float t[64][33];
    const int x0 = blockIdx.x * 64, y = blockIdx.y, z0 = blockIdx.z * 32;
    const float *src = vol + (size_t)y * w * nz;
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int xl = e >> 5, zl = e & 31;
        const int x = x0 + xl, z = z0 + zl;
        t[xl][zl] = (x < w && z < nz) ? src[(size_t)x * nz + z] : 0.0f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int zl = e >> 6, xl = e & 63;
        const int x = x0 + xl, z = z0 + zl;
        if (x < w && z < nz) out[((size_t)z * w + y) * w + x] = t[xl][zl];
    }
}

}  // namespace mr

extern "C" int mr_copy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    if (bytes == 0) return MR_OK;
    MR_REQUIRE(d_dst != nullptr && h_src != nullptr, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (mr::is_pinned(h_src, bytes)) {
        MR_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
        MR_HIP(hipStreamSynchronize(s));
        return MR_OK;
    }
    mr::Staging *st = nullptr;
    int rc = mr::staging(0, &st);
    if (rc != MR_OK) return rc;
    std::lock_guard<std::mutex> lk(st->mu);
    const char *src = static_cast<const char *>(h_src);
    char *dst = static_cast<char *>(d_dst);
    int i = 0;
    for (size_t off = 0; off < bytes; off += mr::kChunk, i ^= 1) {
        const size_t n = std::min(mr::kChunk, bytes - off);
        // the DMA that last read this staging buffer must be done
        if (st->used[i]) MR_HIP(hipEventSynchronize(st->done[i]));
        mr::parallel_memcpy(st->buf[i], src + off, n);
        MR_HIP(hipMemcpyAsync(dst + off, st->buf[i], n, hipMemcpyHostToDevice, s));
        MR_HIP(hipEventRecord(st->done[i], s));
        st->used[i] = true;
    }
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_copy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    if (bytes == 0) return MR_OK;
    MR_REQUIRE(h_dst != nullptr && d_src != nullptr, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (mr::is_pinned(h_dst, bytes)) {
        MR_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
        MR_HIP(hipStreamSynchronize(s));
        return MR_OK;
    }
    mr::Staging *st = nullptr;
    int rc = mr::staging(1, &st);
    if (rc != MR_OK) return rc;
    std::lock_guard<std::mutex> lk(st->mu);
    const char *src = static_cast<const char *>(d_src);
    char *dst = static_cast<char *>(h_dst);
    // chunk k goes device -> buf[k % 2]; while it moves, chunk k - 1 is
    // copied out of the other buffer
    const size_t nchunks = (bytes + mr::kChunk - 1) / mr::kChunk;
    for (size_t k = 0; k <= nchunks; ++k) {
        if (k < nchunks) {
            const int i = (int)(k & 1);
            const size_t off = k * mr::kChunk, n = std::min(mr::kChunk, bytes - off);
            MR_HIP(hipMemcpyAsync(st->buf[i], src + off, n, hipMemcpyDeviceToHost, s));
            MR_HIP(hipEventRecord(st->done[i], s));
            st->used[i] = true;
        }
        if (k >= 1) {
            const int j = (int)((k - 1) & 1);
            const size_t off = (k - 1) * mr::kChunk, n = std::min(mr::kChunk, bytes - off);
            MR_HIP(hipEventSynchronize(st->done[j]));
            mr::parallel_memcpy(dst + off, st->buf[j], n);
        }
    }
    return MR_OK;
    MR_GUARD_END
}

extern "C" int mr_volume_zxy_to_xyz(const float *d_zxy, float *d_xyz, size_t width, size_t slices, void *stream)
{
    MR_GUARD_BEGIN
    mr::clear_error();
    MR_REQUIRE(d_zxy != nullptr && d_xyz != nullptr, "NULL pointer");
    MR_REQUIRE(d_zxy != d_xyz, "the transposition is out of place");
    MR_REQUIRE(width >= 1 && width <= 65536 && slices >= 1 && slices <= 65535 * 32, "size out of range");
    if (width == 0 || slices == 0) return MR_OK;
    hipStream_t s = (hipStream_t)stream;
    const dim3 g((unsigned)((width + 63) / 64), (unsigned)width, (unsigned)((slices + 31) / 32));
    mr::ProfScope prof("zxy_to_xyz", s);
    hipLaunchKernelGGL(mr::zxy_to_xyz_kernel, g, dim3(256), 0, s, d_zxy, d_xyz, (int)width, (int)slices);
    MR_HIP(hipGetLastError());
    return MR_OK;
    MR_GUARD_END
}
