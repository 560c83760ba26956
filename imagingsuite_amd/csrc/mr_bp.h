// Internal interface of the parallel-beam back-projection (mr_bp.hip) for
// the other translation units of the library: the plan's resident-table
// path and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "muhrec_hip.h"

namespace mr {

// Arguments of the back-projection kernels (one stack).
struct BpArgs {
    const float *sino;
    float *vol;
    const float4 *tab;   // (fStartU, fCos, fSin, 0) per used projection
    const int2 *mask;    // [first, second) per y, 0-based x
    int nu, nv, vpad, nk, nz4, bstride, accumulate;
    int order;           // block -> tile mapping (tile_of_block)
    int ntx, nty, ntz;   // tiles along x, y and slice chunks
    int sw;              // stripe width (tiles) of the XCD-striped order
    int stage64;         // 1: stage with 64-bit addresses even when 32-bit offsets fit (tests)
    // axis-tilt branch: 0 none, 1 MultiProjBPparallel, 2 MultiProjBP,
    // 3 NearestNeighborBP; centeroffset and centerinc of the block
    int tilt;
    float tilt_co, tilt_ci;
};

// Uploads the tables of n_proj projections (angles, weights) into the plan
// and sizes its buffers; synchronises s (the host tables are temporaries).
int bp_load_tables(mr_bp_plan *plan, size_t n_proj, const float *angles, const float *weights, hipStream_t s);
// Weighting/transposition and back-projection of d_proj [n_proj][slices][width]
// with the loaded tables into d_volume [width][width][slices]; no host sync.
int bp_enqueue(mr_bp_plan *plan, const float *d_proj, float *d_volume, int accumulate, hipStream_t s);
size_t bp_plan_slices(const mr_bp_plan *plan);
// First detector row of the next bp_enqueue's block (the tilt terms' roi[1]);
// no effect without tilt correction.
void bp_set_tilt_row0(mr_bp_plan *plan, size_t row0);
size_t bp_plan_width(const mr_bp_plan *plan);

// sino[k][u][v] = proj[idx[k]][v][u] * wts[idx[k]] for the nk used
// projections (v < nv; zero for nv <= v < vpad).
int launch_bp_prep(const float *proj, float *sino, const int *idx, const float *wts, int nu, int nv,
                   int vpad, int nk, hipStream_t s);
// Back-projection kernel for one stack.  impl 0: direct gathers (MultiProjBP
// with slices % 4 != 0, whose second column is read at a different offset);
// 1: LDS-staged tiles.  exact: the reference's float operation order.
int run_bp(const BpArgs &a, int interp, int impl, int exact, hipStream_t s);

}  // namespace mr
