// Internal launch interface of the matched Joseph pair (mr_joseph.hip), used
// by the iterative solvers to run A and A^T with resident tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "muhrec_hip.h"

namespace mr {

// Per angle (su, cos, sin, 1/m) as four floats.
std::vector<float> joseph_tables(const mr_joseph_geom &g, const float *angles, size_t n_ang);
int joseph_check(const mr_joseph_geom *g, size_t n_ang);
// Enqueues A (forward: volume -> sinogram) or A^T on s; d_tab holds n_ang
// table entries on the device.  Does not synchronise.
int joseph_launch(const mr_joseph_geom *g, const float *src, float *dst, const float4 *d_tab, size_t n_ang,
                  bool forward, hipStream_t s);

}  // namespace mr
