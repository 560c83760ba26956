/*
 * muhrec_hip.h -- C ABI of the MI355X (gfx950) reconstruction kernels.
 *
 * Plain pointers and sizes only.  Every function returns an mr_status code;
 * on failure mr_last_error() returns a thread-local message.  No C++
 * exception crosses this boundary.  Pointers named d_* are device pointers
 * (hipMalloc'd, or a torch CUDA tensor's data_ptr); everything else is host
 * memory.  `stream` is a hipStream_t (NULL = the default stream).  Plans are
 * bound to the device that is current when they are created.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the DTasev/imagingsuite root).
 */
#ifndef MUHREC_HIP_H
#define MUHREC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mr_status {
    MR_OK = 0,
    MR_ERR_INVALID = 1,     /* bad argument (the reference throws ReconException) */
    MR_ERR_HIP = 2,         /* a HIP runtime call failed */
    MR_ERR_UNSUPPORTED = 3, /* configuration outside what the kernels implement */
    MR_ERR_INTERNAL = 4,    /* unexpected host-side exception */
    MR_ERR_ABORTED = 5      /* the caller asked to stop (a callback returned nonzero) */
};

/* ProjectionFilterBase::FilterType,
 * frameworks/tomography/Preprocessing/StdPreprocModules/include/ProjectionFilter.h:27-33 */
enum mr_filter_type {
    MR_FILTER_RAMLAK = 0,
    MR_FILTER_SHEPPLOGAN = 1,
    MR_FILTER_HANNING = 2,
    MR_FILTER_HAMMING = 3,
    MR_FILTER_BUTTERWORTH = 4
};

enum mr_interp {
    MR_INTERP_LINEAR = 0,  /* MultiProjBP / MultiProjBPparallel */
    MR_INTERP_NEAREST = 1  /* NearestNeighborBP */
};

/* Message of the last failing call on this thread ("" if none). */
const char *mr_last_error(void);
/* ABI version of this library. */
int mr_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int mr_device_count(void);

/* Kernel timing (no reference counterpart; measurement hook for benchmarks).
 * While enabled, every kernel launch of this library is bracketed by HIP
 * events on the stream it is launched on.  mr_profile_query synchronises on
 * the recorded events and returns the summed duration and launch count of
 * the kernels whose name contains `kernel` ("" = all). */
int mr_profile_enable(int on);
int mr_profile_reset(void);
int mr_profile_query(const char *kernel, double *total_ms, long *launches);

/* ------------------------------------------------------------------------
 * ProjectionFilterSingle ramp filter.
 * Replaces ProjectionFilterBase::Configure + ProjectionFilterSingle::BuildFilter
 * (frameworks/tomography/Preprocessing/StdPreprocModules/src/ProjectionFilter.cpp:78-92,254-302)
 * and FilterProjection / ProcessCore(3D) (:103-118, :304-344).
 * ------------------------------------------------------------------------ */
typedef struct mr_filter_plan mr_filter_plan;

/* width = ROI width N; the FFT length is 2^(ceil(log2 N)+padding_doubler).
 * Fails with MR_ERR_INVALID for a cut-off outside (0, 0.5] (reference :265-266). */
int mr_filter_plan_create(mr_filter_plan **plan, size_t width, int filter_type, float cutoff,
                          float order, int use_bias, float bias_weight, size_t padding_doubler);
int mr_filter_plan_destroy(mr_filter_plan *plan);
/* fft_size, n_insert and (optionally) the fft_size/2 filter coefficients H. */
int mr_filter_plan_info(const mr_filter_plan *plan, size_t *fft_size, size_t *n_insert,
                        float *h_filter);
/* The kernel form this plan launches, e.g. "row_filter_kernel2<4096,512> short"
 * (the filter as a 4096-point convolution plus the taper responses, when
 * 2 width - 1 fits a power of two below the FFT length) or
 * "row_filter_kernel2<8192,1024>".  Diagnostics only; no counterpart in the
 * reference. */
int mr_filter_plan_kernel(const mr_filter_plan *plan, char *name, size_t len);
/* Filters n_rows rows of `width` samples in place (device memory).  A projection
 * stack [n_proj][rows][width] is n_proj*rows rows. */
int mr_filter_rows(mr_filter_plan *plan, float *d_rows, size_t n_rows, void *stream);
/* Adjoint of mr_filter_rows, in place (no reference counterpart; for
 * gradients through FBP): the rows are embedded at nInsert, convolved with the
 * same symmetric ramp, cropped, and the two tapers are folded back onto the
 * first and last sample. */
int mr_filter_rows_adjoint(mr_filter_plan *plan, float *d_rows, size_t n_rows, void *stream);

/* ------------------------------------------------------------------------
 * Parallel-beam back-projection.
 * Replaces StdBackProjectorBase::SetROI/Process(2D/3D)
 * (frameworks/tomography/Backprojectors/StdBackProjectors/src/StdBackProjectorBase.cpp:64-201),
 * MultiProjectionBPparallel::BackProject (.../src/MultiProjBPparallel.cpp:39-169),
 * MultiProjectionBP::BackProject (.../src/MultiProjBP.cpp:36-144) and
 * NearestNeighborBP::BackProject (.../src/NNMultiProjBP.cpp:36-119),
 * with BackProjectorModuleBase::BuildCircleMask
 * (frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp:153-218).
 * ------------------------------------------------------------------------ */
typedef struct {
    size_t width;        /* SizeU = roi[2]-roi[0]; volume is width x width x slices */
    size_t slices;       /* SizeV = detector rows of the block */
    float center;        /* ProjectionInfo.fCenter (ROI-relative) */
    int direction;       /* kipl::base::eRotationDirection: 0 CW, 1 CCW */
    int interp;          /* mr_interp */
    int buffer_size;     /* ProjectionBufferSize; >0 reproduces the reference's
                            drop of every buffer_size-th projection, 0 uses all */
    int multiprojbp;     /* 1: MultiProjBP column-B offset (differs only when slices%4 != 0) */
    int use_roi;         /* MatrixInfo.bUseROI */
    size_t roi[4];       /* MatrixInfo.roi x0,y0,x1,y1 (inclusive) */
    int fused;           /* 0: the reference's float operation order per voxel (bit-identical
                            results); 1: interpolation folded into fused multiply-adds
                            (differs by float rounding only) */
    /* Axis-tilt correction, the bCorrectTilt branches of the three back-projectors
     * (MultiProjBPparallel.cpp:49-114, MultiProjBP.cpp:50-96, NNMultiProjBP.cpp:48-81)
     * and the mask shrink of BuildCircleMask (BackProjectorModuleBase.cpp:171-180).
     * mr_bp_parallel, mr_fbp_parallel and the mr_fbp_stream entry points honour it
     * (always in the reference's float order: `fused` is ignored); the adjoint and
     * the no-allocation entry point refuse it. */
    int correct_tilt;    /* ProjectionInfo.bCorrectTilt */
    float tilt_angle;    /* ProjectionInfo.fTiltAngle (degrees) */
    float tilt_pivot;    /* ProjectionInfo.fTiltPivotPosition (detector row) */
    size_t tilt_row0;    /* ProjectionInfo.roi[1] of this block: its first detector row
                            (the stream entry points add each slab's offset) */
    size_t tilt_rows;    /* ProjectionInfo.nDims[1] (roi[3] for repeated sinograms):
                            the mask radius shrinks by floor(tan|angle| * tilt_rows) */
} mr_bp_params;

typedef struct mr_bp_plan mr_bp_plan;

int mr_bp_plan_create(mr_bp_plan **plan, const mr_bp_params *params);
int mr_bp_plan_destroy(mr_bp_plan *plan);
/* Name of the kernel variant mr_bp_parallel launches for this plan (slice
 * tile, projection chunk, block order), e.g.
 * "bp_lds_kernel<32,0,false,4> order=1 sw=4".  Diagnostics only; no
 * counterpart in the reference. */
int mr_bp_plan_kernel(const mr_bp_plan *plan, char *name, size_t len);
/* Host copy of the circle mask: for row y the voxels first[y] <= x < second[y]
 * are reconstructed (BuildCircleMask's pair, with `second` exclusive). */
int mr_bp_plan_mask(const mr_bp_plan *plan, size_t *first, size_t *second);
/* Back-projects n_proj filtered projections d_proj [n_proj][slices][width]
 * (the reference's TImage<float,3> layout) with host angles (degrees) and
 * weights into d_volume [width(y)][width(x)][slices] (MatrixZXY).
 * accumulate = 0 overwrites the volume (SetROI zeroes it); 1 adds to it.
 * The kernels are enqueued on `stream` and the call returns without waiting
 * for them (angles and weights are copied before it returns); the next call
 * on the same plan waits for them on the device, whatever its stream.
 * mr_fbp_parallel behaves the same way. */
int mr_bp_parallel(mr_bp_plan *plan, const float *d_proj, size_t n_proj, const float *angles,
                   const float *weights, float *d_volume, int accumulate, void *stream);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  mr_bp_parallel without a plan and without any allocation:
 * the circle mask, the projection tables and the weighted, transposed
 * sinogram live in a fixed device array of MR_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  The used projections are back-projected in
 * chunks of as many as fit in that array; the result equals mr_bp_parallel's
 * for the same params bit for bit.  Fails with MR_ERR_INVALID when one
 * projection's staged size (width x slices rounded up to 4, float) exceeds
 * the array's 60 MiB sinogram part, and for more than 65535 projections.
 * Enqueues on `stream` (of the current device) and returns without waiting.
 * All calls on one GPU share the array: calls that overlap on one GPU
 * (other streams, other threads) race, and must be ordered by the caller. */
#define MR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int mr_bp_parallel_noalloc(const mr_bp_params *params, const float *d_proj, size_t n_proj, const float *angles,
                           const float *weights, float *d_volume, int accumulate, void *stream);
/* Adjoint of mr_bp_parallel (the project's own operator for gradients; no
 * reference counterpart): d_grad_volume [width(y)][width(x)][slices] ->
 * d_grad_proj [n_proj][slices][width] (overwritten; zero for dropped
 * projections and slices >= 4*(slices/4)).  Uses the forward's float
 * positions exactly.  Not available for MultiProjBP with slices % 4 != 0. */
int mr_bp_parallel_adjoint(mr_bp_plan *plan, const float *d_grad_volume, size_t n_proj, const float *angles,
                           const float *weights, float *d_grad_proj, void *stream);
/* Filter (in place on d_proj) followed by back-projection: one engine block of
 * ReconEngine::Process3D (frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:1499-1511,1663-1714). */
int mr_fbp_parallel(mr_filter_plan *fplan, mr_bp_plan *bplan, float *d_proj, size_t n_proj,
                    const float *angles, const float *weights, float *d_volume, void *stream);

/* ------------------------------------------------------------------------
 * FullLogNorm (the normalisation in front of the filter), non-LUT path.
 * Replaces FullLogNorm::SetReferenceImages and ProcessCore(3D)
 * (frameworks/tomography/Preprocessing/StdPreprocModules/src/NormPlugins.cpp:300-341, 411-558).
 * Reading and averaging the reference images (NormBase::ReferenceLoader) is
 * file I/O and stays with the caller.
 * ------------------------------------------------------------------------ */
/* d_flat (device, n_pix) becomes log((flat - dark) / flat_dose), 0 where
 * flat - dark <= 0.  d_dark may be NULL (no dark current images). */
int mr_lognorm_reference(float *d_flat, const float *d_dark, size_t n_pix, float flat_dose, void *stream);
/* In place on d_proj [n_proj][n_pix]: p = img - dark (or img),
 * img = p <= 0 ? 0 : flat_log - log(p) + log(max(dose[j] - dark_dose, 1)).
 * d_flat_log / d_dark may be NULL (nOBCount / nDCCount == 0); dose (host,
 * n_proj values) may be NULL (usenormregion = false: dose term 0). */
int mr_lognorm(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
               const float *dose, float dark_dose, void *stream);
/* The same with the per-projection dose term given directly (host, n_proj
 * values, or NULL for 0): the 2-D path's log(dose <= 0 ? 1 : dose)
 * (ProcessCore(2D), NormPlugins.cpp:343-357). */
int mr_lognorm_logdose(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                       const float *dose_log, void *stream);
/* FullLogNorm's look-up-table path (uselut=true; NormPlugins.cpp:455-475 with
 * kipl's LogLUT(65536, 0, 65535), LUTCollection.cpp:11-36, LUTbase.cpp:60-78):
 * p = img - dark (or img); L = log(1) = 0 for p < 0, log(65535) for
 * p > 65535, else log of the entry (size_t)(p + 0.5), entry 0 being log(1);
 * img = flat_log - L + doselist[j], or -L + doselist[j] - doselist[0] without
 * a flat (d_flat_log NULL), doselist as in mr_lognorm.  Unlike the non-LUT
 * path there is no p <= 0 -> 0 case.  The table is computed on the host and
 * held once per GPU. */
int mr_lognorm_lut(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                   const float *dose, float dark_dose, void *stream);
/* The same with each projection's additive term given (host, n_proj values, or
 * NULL for 0): the 2-D path's cLogDose (NormPlugins.cpp:370-381). */
int mr_lognorm_lut_term(float *d_proj, size_t n_proj, size_t n_pix, const float *d_flat_log, const float *d_dark,
                        const float *term, void *stream);
/* FullLogNorm fused into ProjectionFilterSingle's row load: d_proj
 * [n_proj][rows][width] raw counts become the filtered -log projections in
 * one pass (arguments as mr_lognorm, n_pix = rows * width); the same values
 * as mr_lognorm followed by mr_filter_rows.  No reference counterpart: the
 * reference runs the two modules one after the other. */
int mr_filter_rows_lognorm(mr_filter_plan *plan, float *d_proj, size_t n_proj, size_t rows,
                           const float *d_flat_log, const float *d_dark, const float *dose, float dark_dose,
                           void *stream);

/* ------------------------------------------------------------------------
 * Streaming FBP of a host-resident stack (engine side).  Replaces the
 * SliceBlock loop of ReconEngine::Run3DFull/Process3D
 * (frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:838-1185,
 * 1314-1714) for the ProjectionFilterSingle + parallel back-projector chain,
 * and its slice transfer (TransferMatrix, :531-574): each block of
 * slice_block detector rows is copied to the GPU, filtered, weighted and
 * back-projected there, transposed to slices and copied back, with the
 * copies of neighbouring blocks overlapped with the kernels.  The results
 * equal, bit for bit, mr_filter_rows + mr_bp_parallel run block by block.
 * bp->slices is ignored (the block height is used).  h_proj:
 * [n_proj][rows][width]; h_volume: [rows][width][width] (slice, y, x: the
 * order of GetSlice).  Host memory from mr_host_alloc is transferred by DMA
 * directly; other host memory is staged through pinned buffers.
 * serial = 1 waits for each block before starting the next (no overlap; a
 * baseline for timing).
 * ------------------------------------------------------------------------ */
typedef struct mr_fbp_stream mr_fbp_stream;
int mr_fbp_stream_create(mr_fbp_stream **stream, const mr_bp_params *bp, size_t rows, size_t n_proj,
                         size_t slice_block, int filter_type, float cutoff, float order, int use_bias,
                         float bias_weight, size_t padding_doubler);
int mr_fbp_stream_run(mr_fbp_stream *stream, const float *h_proj, const float *angles, const float *weights,
                      float *h_volume, int serial);
/* The same run with each slab handed over as soon as its slices are in
 * h_volume, while the GPU works on the later slabs: the engine's per-block
 * output (ReconEngine::Process3D's TransferMatrix and, with
 * bAutomaticSerialize, the write of the block's slices,
 * frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:1703-1710)
 * runs beside the reconstruction instead of after it.  on_slab(user,
 * first_slice, n_slices, slices) is called on the calling thread, slabs in
 * order; `slices` points into h_volume ([n_slices][width][width]).  A nonzero
 * return stops the run (MR_ERR_ABORTED; the later slabs are not delivered). */
typedef int (*mr_slab_fn)(void *user, size_t first_slice, size_t n_slices, const float *slices);
int mr_fbp_stream_run_slabs(mr_fbp_stream *stream, const float *h_proj, const float *angles, const float *weights,
                            float *h_volume, mr_slab_fn on_slab, void *user);
int mr_fbp_stream_destroy(mr_fbp_stream *stream);
/* Pinned (page-locked) host memory that the streaming path copies by DMA. */
int mr_host_alloc(void **ptr, size_t bytes);
int mr_host_free(void *ptr);

/* Copies between host memory and the device for callers that hold pageable
 * host buffers, as MuhRec's modules do (kipl TImage; no reference
 * counterpart: the reference has no device).  Pageable memory goes through
 * two pinned staging buffers of the library (64 MiB each per device, kept
 * until the process ends), filled or drained by host threads while the DMA
 * engine moves the other one; mr_host_alloc memory is copied by DMA
 * directly.  Both return when the copy is complete; calls from several
 * threads on one device take turns. */
int mr_copy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int mr_copy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
/* d_xyz[z][y][x] = d_zxy[y][x][z] for a width x width x slices volume: the
 * back-projector's MatrixZXY volume to MatrixXYZ, whose slices
 * BackProjectorModuleBase::GetSlice copies row by row
 * (frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp:62-124).
 * Out of place; enqueued on `stream` without waiting. */
int mr_volume_zxy_to_xyz(const float *d_zxy, float *d_xyz, size_t width, size_t slices, void *stream);

/* ------------------------------------------------------------------------
 * FDK cone-beam back-projection (FDKbp_single and FDKbp).
 * Replaces FdkReconBase::SetROI/Process(3D)
 * (frameworks/tomography/Backprojectors/FDKBackProjectors/src/fdkreconbase.cpp:91-277),
 * FDKbp_single::getProjMatrix / ramp_filter_tuned / project_volume_onto_image_c
 * (.../src/fdkbp_single.cpp:83-255, 832-958, 293-531, 641-683) and the
 * double-geometry FDKbp versions of the same three (.../src/fdkbp.cpp:83-257,
 * 822-948, 324-510, 612-631).
 * ------------------------------------------------------------------------ */
#define MR_FDK_SINGLE 0  /* FDKbp_single: float geometry, scale SOD/SDD */
#define MR_FDK_DOUBLE 1  /* FDKbp: double geometry and sampling, scale SOD/SDD/nProj */
typedef struct {
    float sod, sdd;         /* ProjectionInfo.fSOD, fSDD (mm) */
    float pixel[2];         /* ProjectionInfo.fResolution (mm) */
    float voxel[3];         /* MatrixInfo.fVoxelSize (mm) */
    float ppoint[2];        /* ProjectionInfo.fpPoint (piercing point, detector px) */
    float center;           /* ProjectionInfo.fCenter */
    int direction;          /* 0 CW, 1 CCW */
    size_t roi[4];          /* ProjectionInfo.roi of this slab (x0,y0,x1,y1) */
    size_t projection_roi3; /* ProjectionInfo.projection_roi[3] */
    int use_roi;            /* MatrixInfo.bUseROI */
    size_t matrix_roi[4];   /* MatrixInfo.roi */
    int module;             /* MR_FDK_SINGLE or MR_FDK_DOUBLE */
    int fused;              /* MR_FDK_SINGLE only.  0: the reference's float operations per sample
                               (bit-identical positions); 1: row position and accumulation folded
                               into FMAs (a position on a pixel edge may pick the neighbour) */
    /* Filter extension without reference counterpart (both modules; 0/0 is the reference):
     * filter 0 = ramp_filter_tuned's ramp and roll-off, 1 = Shepp-Logan |k|/(P/2) sinc(k/P);
     * cosine_weight 1 multiplies each detector pixel by SDD / sqrt(SDD^2 + u^2 + v^2)
     * before the filter (u, v in mm from the piercing point). */
    int filter;
    int cosine_weight;
    /* Axis-tilt correction (fdkbp_single.cpp:443-448, fdkbp.cpp:452-458): slice k's
     * centre of rotation tan(-angle) (window row3 - k - pivot) + center replaces the
     * projection matrix's column offset; the mask radius shrinks by
     * floor(tan|angle| * tilt_rows) (BackProjectorModuleBase.cpp:171-180). */
    int correct_tilt;       /* ProjectionInfo.bCorrectTilt */
    float tilt_angle;       /* ProjectionInfo.fTiltAngle (degrees) */
    float tilt_pivot;       /* ProjectionInfo.fTiltPivotPosition */
    size_t tilt_rows;       /* ProjectionInfo.nDims[1] (roi[3] for repeated sinograms) */
} mr_fdk_params;

typedef struct mr_fdk_plan mr_fdk_plan;

/* Host-only: the detector-row window (CBCT_roi x0,y0,x1,y1) of the slab in
 * params->roi, as ReconEngine::Run3DFull computes it
 * (frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:954-1008). */
int mr_fdk_window(const mr_fdk_params *params, size_t *cbct_roi);
int mr_fdk_plan_create(mr_fdk_plan **plan, const mr_fdk_params *params);
int mr_fdk_plan_destroy(mr_fdk_plan *plan);
/* Detector window (CBCT_roi x0,y0,x1,y1) whose rows the caller hands to
 * mr_fdk_bp, as ReconEngine::Run3DFull computes it (ReconEngine.cpp:954-1008). */
int mr_fdk_plan_window(const mr_fdk_plan *plan, size_t *cbct_roi);
/* d_proj: [n_proj][win_rows][width] raw projections of the window (device,
 * used as scratch: it is weighted and filtered in place).  d_volume:
 * [nz][nx][nx] (MatrixXYZ, x fastest) with nx = roi[2]-roi[0], nz = roi[3]-roi[1];
 * overwritten with the mirrored/flipped result (fdkreconbase.cpp:259-270). */
int mr_fdk_bp(mr_fdk_plan *plan, float *d_proj, size_t n_proj, size_t win_rows,
              const float *angles, const float *weights, float *d_volume, void *stream);

/* ------------------------------------------------------------------------
 * Batched reconstruction of many parallel-beam stacks of uneven geometry
 * (no reference counterpart: the reference reconstructs one stack per engine
 * run, ReconEngine::Run3DFull, frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:838-1185).
 * Each stack is filtered (ProjectionFilterSingle, out of place: d_proj is not
 * modified), weighted and back-projected (MultiProjBPparallel order of
 * operations, bit-identical per stack).  mr_batch_run enqueues kernels only
 * and ends with ONE persistent back-projection launch over a work queue of
 * all stacks' tiles; mr_batch_run_graph captures it into a hipGraph on the
 * first call and replays it after; mr_batch_run_sequential launches one
 * plain back-projection per stack (the comparison baseline).
 * Device buffers are bound at plan creation.
 * ------------------------------------------------------------------------ */
typedef struct {
    size_t width, slices, n_proj;   /* detector width, rows, projections */
    float center;                   /* fCenter (ROI-relative) */
    int direction;                  /* 0 CW, 1 CCW */
    int filter_type;                /* mr_filter_type */
    float cutoff, order;
    int use_bias;
    float bias_weight;
    size_t padding_doubler;
    int buffer_size;                /* ProjectionBufferSize drop (0 = all) */
    const float *d_proj;            /* [n_proj][slices][width] raw projections */
    const float *angles, *weights;  /* host, n_proj each */
    float *d_volume;                /* [width][width][slices] */
} mr_batch_stack;

typedef struct mr_batch_plan mr_batch_plan;

/* flags: MR_BATCH_TUNE - the autotuner times every (slice tile, projection
 * chunk) variant of the persistent kernel on this batch and keeps the
 * fastest; MR_BATCH_FUSED - the interpolation is folded into two FMAs per
 * update (as mr_bp_params.fused; float rounding only), otherwise every stack
 * keeps MultiProjBPparallel's operation order.  Unknown bits: MR_ERR_INVALID. */
#define MR_BATCH_TUNE 1
#define MR_BATCH_FUSED 2
int mr_batch_plan_create(mr_batch_plan **plan, const mr_batch_stack *stacks, size_t n_stacks, int flags);
int mr_batch_plan_info(const mr_batch_plan *plan, int *slice_tile, int *proj_chunk, size_t *n_items,
                       double *updates);
int mr_batch_run(mr_batch_plan *plan, void *stream);
int mr_batch_run_graph(mr_batch_plan *plan, void *stream);
int mr_batch_run_sequential(mr_batch_plan *plan, void *stream);
int mr_batch_plan_destroy(mr_batch_plan *plan);

/* ------------------------------------------------------------------------
 * Matched Joseph projector pair for gradient-based reconstruction.
 * No reference counterpart: the reference's only forward projector,
 * LinearForwardProjector::project
 * (frameworks/tomography/Framework/ReconAlgorithms/ReconAlgorithms/linearforwardprojector.cpp:15-143),
 * is 2-D and not the adjoint of any back-projector.  The pair uses the
 * parallel-beam geometry of StdBackProjectorBase (.../StdBackProjectorBase.cpp:73-80)
 * and mr_joseph_adjoint is the exact transpose of mr_joseph_forward.
 * volume [nx(y)][nx(x)][nz] (MatrixZXY), sinogram [n_ang][n_det][nz]; nz % 4 == 0.
 * ------------------------------------------------------------------------ */
typedef struct {
    size_t nx;      /* volume is nx x nx x nz */
    size_t nz;      /* slices (multiple of 4) */
    size_t n_det;   /* detector columns */
    float center;   /* detector column of the rotation axis (fCenter) */
    int direction;  /* 0 CW, 1 CCW */
} mr_joseph_geom;

int mr_joseph_forward(const mr_joseph_geom *geom, const float *d_volume, const float *angles,
                      size_t n_ang, float *d_sino, void *stream);
int mr_joseph_adjoint(const mr_joseph_geom *geom, const float *d_sino, const float *angles,
                      size_t n_ang, float *d_volume, void *stream);

/* LinearForwardProjector::project
 * (frameworks/tomography/Framework/ReconAlgorithms/ReconAlgorithms/linearforwardprojector.cpp:15-143)
 * for a stack of slices: d_volume [M(y)][M(x)][nz] (MatrixZXY; nz = 1 is the
 * reference's single TImage<float,2> slice, slice(x, y) = d_volume[y M + x]),
 * d_sino [n_ang][M][nz] (proj(r, t) of slice z at [(t M + r) nz + z]).  Each
 * slice's projection equals the reference's bit for bit, including its
 * quirks (bins whose position starts below 0 at 90 degrees lose the steps
 * before it turns non-negative, the last bin of an axis-aligned view is 0).
 * The reference throws for 3-D input; here slices are independent. */
int mr_linear_forward(const float *d_volume, size_t M, size_t nz, const float *angles, size_t n_ang,
                      float *d_sino, void *stream);

/* ------------------------------------------------------------------------
 * Iterative reconstruction on the matched pair: SIRT, Landweber and CGLS as
 * sketched in frameworks/tomography/Backprojectors/IterativeBackProj/src/IterativeMeth.txt
 * (landweberOTF :5-17, SIRT_OTF :21-52, cgLS_OTF :56-88); replaces the
 * reference's SIRTbp::Process (.../src/sirtbp.cpp:32-81), which is a stub.
 * The plan holds the angle table and work buffers (SIRT: R = 1/A1 and
 * C = 1/A^T 1, computed at creation).  mr_iter_run iterates on d_volume
 * (initial guess in, result out; [nx][nx][nz]) against d_sino
 * ([n_ang][n_det][nz]); it enqueues on `stream` without host synchronisation
 * unless `residuals` (host, `iterations` values of |b - A x|^2 before each
 * update) is requested.  tau: Landweber step (> 0; ignored otherwise);
 * nonneg: -1 = method default (Landweber clamps at 0 like landweberOTF,
 * SIRT does not), 0/1 = off/on (CGLS: off only).  Calls on one plan must not
 * overlap.
 * ------------------------------------------------------------------------ */
#define MR_ITER_SIRT 0
#define MR_ITER_LANDWEBER 1
#define MR_ITER_CGLS 2
typedef struct mr_iter_plan mr_iter_plan;
int mr_iter_plan_create(mr_iter_plan **plan, const mr_joseph_geom *geom, const float *angles, size_t n_ang,
                        int method);
int mr_iter_run(mr_iter_plan *plan, const float *d_sino, float *d_volume, int iterations, float tau, int nonneg,
                double *residuals, void *stream);
int mr_iter_plan_destroy(mr_iter_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* MUHREC_HIP_H */
