// Error state, version query and small utility kernels of the C ABI.
#include "mr_common.h"
#include "mr_filter.h"

#include <atomic>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

namespace mr {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
    return MR_ERR_HIP;
}

namespace {
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof;
std::atomic<bool> g_prof_on{false};
}  // namespace

ProfScope::ProfScope(const char *name, hipStream_t s) : slot_(-1), s_(s)
{
    if (!g_prof_on.load()) return;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return;
    ProfRec r{name, nullptr, nullptr};
    if (hipEventCreate(&r.a) != hipSuccess) return;
    if (hipEventCreate(&r.b) != hipSuccess) { (void)hipEventDestroy(r.a); return; }
    if (hipEventRecord(r.a, s) != hipSuccess) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        return;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    slot_ = (long)g_prof.size();
    g_prof.push_back(r);
}

ProfScope::~ProfScope()
{
    if (slot_ < 0) return;
    hipEvent_t b;
    {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        if (slot_ >= (long)g_prof.size()) return;
        b = g_prof[slot_].b;
    }
    (void)hipEventRecord(b, s_);
}

__global__ void fill_zero_kernel(float *d, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = 0.0f;
}

int fill_zero(float *d, size_t n, hipStream_t s)
{
    if (n == 0) return MR_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    ProfScope prof("fill_zero", s);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr

extern "C" const char *mr_last_error(void) { return mr::g_last_error.c_str(); }

extern "C" int mr_version(void) { return 1; }

extern "C" int mr_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int mr_profile_enable(int on)
{
    mr::g_prof_on.store(on != 0);
    return MR_OK;
}

extern "C" int mr_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(mr::g_prof_mu);
    for (auto &r : mr::g_prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    mr::g_prof.clear();
    return MR_OK;
}

extern "C" int mr_profile_query(const char *kernel, double *total_ms, long *launches)
{
    mr::clear_error();
    if (!total_ms || !launches) {
        mr::set_error("NULL argument");
        return MR_ERR_INVALID;
    }
    std::string key = kernel ? kernel : "";
    double tot = 0.0;
    long cnt = 0;
    std::lock_guard<std::mutex> lk(mr::g_prof_mu);
    for (auto &r : mr::g_prof) {
        if (!key.empty() && r.name.find(key) == std::string::npos) continue;
        hipError_t e = hipEventSynchronize(r.b);
        if (e != hipSuccess) return mr::hip_fail(e, "hipEventSynchronize", __FILE__, __LINE__);
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e != hipSuccess) return mr::hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        tot += ms;
        ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
    return MR_OK;
}
