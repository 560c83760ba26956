// Host-side geometry and filter set-up shared by the kernels' C ABI.
// Single-precision formulas follow the reference's float operation order so
// that the tables the kernels consume are bit-identical to the values the
// reference computes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace mr {

// ProjectionFilterBase::ComputeFilterSize + ProjectionFilterSingle::BuildFilter
// + PreparePadding (frameworks/tomography/Preprocessing/StdPreprocModules/src/
// ProjectionFilter.cpp:120-125, 254-302, 363-375).  Returns an mr_status.
int build_pb_filter(size_t width, int filter_type, float cutoff, float order, bool use_bias,
                    float bias_weight, size_t padding_doubler, size_t &P, size_t &n_insert,
                    std::vector<float> &H, std::vector<float> &taper);

// BackProjectorModuleBase::BuildCircleMask
// (frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp:153-218).
// first/second are the reference's pair (second inclusive).  shrink is the
// tilt correction's radius reduction (tilt_mask_shrink), 0 without tilt.
void circle_mask(size_t nx, size_t ny, bool use_roi, const size_t *roi, bool cone,
                 std::vector<size_t> &first, std::vector<size_t> &second, float shrink = 0.0f);

// With bCorrectTilt the mask radius shrinks by floor(tan(|angle|) * rows)
// (BackProjectorModuleBase.cpp:171-180; rows = ProjectionInfo.nDims[1], or
// roi[3] for repeated sinograms), in float.
float tilt_mask_shrink(float tilt_angle, size_t rows);

// Tilt terms of the parallel-beam back-projectors' bCorrectTilt branches
// (MultiProjBPparallel.cpp:68-75, MultiProjBP.cpp:63-64, NNMultiProjBP.cpp:61-62):
// centeroffset = (row0 - pivot) tan(angle pi / 180), centerinc = 4 tan(...),
// in float; row0 = ProjectionInfo.roi[1] of the block.
void bp_tilt_terms(float tilt_angle, float tilt_pivot, size_t row0, float &centeroffset, float &centerinc);

// Projections the reference back-projects (StdBackProjectorBase::Process,
// frameworks/tomography/Backprojectors/StdBackProjectors/src/StdBackProjectorBase.cpp:104-114):
// with a buffer of B slots the B-th projection of every chunk is dropped.
std::vector<size_t> used_projections(size_t n_proj, size_t buffer_size);

// Per-projection (fStartU, fCos, fSin) as StdBackProjectorBase.cpp:73-80.
void bp_tables(const float *angles, const size_t *idx, size_t n, int direction, float center,
               size_t matrix_center_x, std::vector<float> &tab4 /* 4 floats per entry */);

struct FdkGeometry {
    float sod, sdd, pixel[2], voxel[3], ppoint[2], center;
    int direction;
    size_t roi[4];
    size_t projection_roi3;
};

// FDKbp_single::getProjMatrix (frameworks/tomography/Backprojectors/
// FDKBackProjectors/src/fdkbp_single.cpp:83-255).
void fdk_proj_matrix(const FdkGeometry &g, float angle, float pm[12]);
// FDKbp::getProjMatrix (fdkbp.cpp:83-257): double vectors and matrices, the
// product accumulated in a float, so every entry is a float value.
void fdkd_proj_matrix(const FdkGeometry &g, float angle, double pm[12]);
// CBCT_roi window (fdkbp_single.cpp:322-367).  fdkbp_variant: FDKbp's copy
// (fdkbp.cpp:361-406), whose "CBCT_roi[1]-8>=0" test on a size_t always
// passes, so a window starting at row 0 wraps to 2^64-8.
void fdk_cbct_roi(const FdkGeometry &g, size_t nvol, size_t cb[4], bool fdkbp_variant = false);
// prepareFFT pad width (fdkbp_single.cpp:960-963).
size_t fdk_padwidth(size_t width);

}  // namespace mr
