// Internal interface of the parallel-beam back-projection plan (mr_bp.hip)
// for callers that keep the projection tables resident across launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "muhrec_hip.h"

namespace mr {

// Uploads the tables of n_proj projections (angles, weights) into the plan
// and sizes its buffers; synchronises s (the host tables are temporaries).
int bp_load_tables(mr_bp_plan *plan, size_t n_proj, const float *angles, const float *weights, hipStream_t s);
// Weighting/transposition and back-projection of d_proj [n_proj][slices][width]
// with the loaded tables into d_volume [width][width][slices]; no host sync.
int bp_enqueue(mr_bp_plan *plan, const float *d_proj, float *d_volume, int accumulate, hipStream_t s);
size_t bp_plan_slices(const mr_bp_plan *plan);
size_t bp_plan_width(const mr_bp_plan *plan);

}  // namespace mr
