// Shared host/device helpers of the MuhRec HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "muhrec_hip.h"

namespace mr {

// Thread-local last error, read through mr_last_error().
void set_error(const std::string &msg);
void clear_error();

// Converts a HIP error into MR_ERR_HIP with a message naming the call site.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Brackets one kernel launch with HIP events on its stream while profiling is
// enabled (mr_profile_enable); a no-op otherwise.
class ProfScope {
public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

private:
    long slot_;
    hipStream_t s_;
};

// Table entry read through the constant address space: with a block-uniform
// index this is a scalar load, so per-projection geometry costs no vector
// load and wait inside the kernels' projection loops.
__device__ __forceinline__ float4 tab_entry(const float4 *tab, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) float4 *)tab)[i];
#else
    return tab[i];   // host pass of a device-only function: never executed
#endif
}

}  // namespace mr

#define MR_HIP(call)                                                              \
    do {                                                                          \
        hipError_t mr_e_ = (call);                                                \
        if (mr_e_ != hipSuccess) return mr::hip_fail(mr_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define MR_REQUIRE(cond, msg)                          \
    do {                                               \
        if (!(cond)) {                                 \
            mr::set_error(std::string(msg));           \
            return MR_ERR_INVALID;                     \
        }                                              \
    } while (0)

// Every entry point is wrapped so that no C++ exception crosses the C ABI.
#define MR_GUARD_BEGIN try {
#define MR_GUARD_END                                           \
    }                                                          \
    catch (const std::exception &ex) {                         \
        mr::set_error(std::string("exception: ") + ex.what()); \
        return MR_ERR_INTERNAL;                                \
    }                                                          \
    catch (...) {                                              \
        mr::set_error("unknown exception");                    \
        return MR_ERR_INTERNAL;                                \
    }
