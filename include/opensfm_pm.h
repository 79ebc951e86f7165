/*
 * opensfm_pm.h -- C ABI of the MI355X PatchMatch multi-view depth estimator.
 *
 * Plain C types only (no torch, no HIP types), so any FFI can bind it.  All
 * functions return PM_OK (0) or a negative PM_E* code; the message of the
 * last failure on the calling thread is available from pm_last_error().
 * One handle = one reference view + its sources + its own HIP stream; a
 * handle is not thread-safe, distinct handles may be used from distinct
 * threads.
 *
 * Reference interfaces replaced (paths relative to the OpenSfM tree):
 *   pm_create / pm_destroy      dense::DepthmapEstimator ctor/dtor
 *                               (opensfm/src/lib/dense/src/estimator.cc:30-31)
 *                               and DepthmapEstimator() of pydense.pyi:42-43
 *   pm_add_view                 DepthmapEstimator::AddView (estimator.cc:33-42),
 *                               add_view(K, R, t, image, mask) (pydense.pyi:45-52)
 *   pm_set_depth_range          SetDepthRange (estimator.cc:44-47),
 *                               set_depth_range(min, max, num_planes) (pydense.pyi:56)
 *   pm_default_params           DepthmapParams{} (estimator.h:16-36)
 *   pm_pipeline_params          the parameters _setup_cluster_params sets
 *                               (opensfm/dense/depthmaps.py:154-175, config.py:407-466)
 *   pm_set_params/pm_get_params SetParams(DepthmapParams) (estimator.cc:49-51,
 *                               estimator.h:16-36); set_patchmatch_iterations,
 *                               set_patch_size, set_min_patch_sd (pydense.pyi:57-59)
 *   pm_set_geom_consistency     SetGeomConsistencyWeight / SetPreviousDepths
 *                               (estimator.cc:53-116)
 *   pm_set_prior                the prior_planes / plane_masks buffers that
 *                               ComputeSfMPlanarPrior uploads (estimator.cc:838-1107)
 *   pm_compute                  Run(DepthmapResult*) (estimator.cc:134-255 with
 *                               RunLevel, BuildImagePyramid, UpsampleDepthNormal +
 *                               ReadBackResults :1112-1224), compute_patch_match()
 *                               (pydense.pyi:54)
 *   pm_upload / pm_run /        the same, split so the device-resident part can be
 *   pm_download                 timed and re-run without host transfers
 *   pm_init_state / pm_iterate  RandomInit + PriorReinit (estimator.cc:515-614) and
 *                               RunIteration (estimator.cc:619-680), for staged parity
 *   pm_read_state/pm_write_state read/write the per-pixel state buffers
 *                               (estimator.cc:452-477)
 *   pm_score_planes             compute_cost_vector + compute_initial_cost_and_views
 *                               (opencl_kernels.h:381-621), the "ComputePlaneScore"
 *                               of the estimator, on caller-given planes
 *   pm_device_count/pm_device_name
 *                               DepthmapClusterEstimator.num_devices / device_name
 *                               (opensfm/src/lib/dense/python/pybind.cc:105-122)
 */
#ifndef OPENSFM_PM_H_
#define OPENSFM_PM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_OK 0
#define PM_E_INVALID (-1)   /* bad argument or call order (std::runtime_error in the reference) */
#define PM_E_DEVICE (-2)    /* HIP runtime failure / no device */
#define PM_E_STATE (-3)     /* operation needs a state that does not exist yet */
#define PM_E_NOMEM (-4)     /* allocation failed */

#define PM_MAX_SOURCES 16   /* reference kMaxSources is 8 (opencl_kernels.h:20) */

/* Pixel formats accepted by pm_add_view. */
#define PM_U8 0     /* 0..255, stored as value/255 (estimator.cc:313-318) */
#define PM_F32 1    /* already in [0, 1] */
#define PM_F16 2    /* IEEE binary16 in [0, 1] */

/* Storage of the images in device memory. */
#define PM_STORE_F32 0
#define PM_STORE_F16 1

typedef struct pm_handle pm_handle;

/* DepthmapParams (estimator.h:16-36) plus the textureless-variance gate
 * (1.5e-6 at opencl_kernels.h:254,516,519) that set_min_patch_sd drives. */
typedef struct pm_params {
  int max_iterations;       /* 3 */
  int patch_size;           /* odd, 3..7 */
  float sigma_spatial;
  float sigma_color;        /* in [0,1] intensity units */
  int top_k;
  int use_census;           /* 0/1 */
  float depth_min, depth_max;
  int hierarchy_levels;     /* 1..8 pyramid levels, coarse to fine (estimator.cc:134-255);
                               level l of N runs at 1/2^(N-1-l) of the input size */
  float smooth_weight;
  int anchor_views;
  int checkerboard_filter;  /* 0/1 */
  float var_threshold;      /* weighted patch variance gate, [0,1]^2 units */
  int image_storage;        /* PM_STORE_F32 / PM_STORE_F16 */
  int quantize_half;        /* 1: round inputs to binary16 like the reference upload */
  int speckle_min_size;     /* > 0: drop connected depth segments smaller than this
                               (RemoveSmallSegments, estimator.cc:1231-1312); 0 = off */
  int gap_max_size;         /* > 0: fill row/column depth gaps up to this length
                               (GapInterpolation, estimator.cc:1319-1418); 0 = off */
} pm_params;

/* The member initialisers of DepthmapParams (estimator.h:16-36): iterations
 * 3, patch 5, sigma_spatial 3, sigma_color 3/255, top_k 4, census on,
 * depth [0, 1], 1 level, smooth 0, anchors 0, filter off, speckle 100,
 * gap 7; plus var_threshold 1.5e-6 (the kernel's gate), f32 storage, no
 * half rounding.  pm_create starts every handle from these, as the
 * reference's DepthmapEstimator starts from a default DepthmapParams. */
void pm_default_params(pm_params* out);
/* The dense pipeline's profile: what _setup_cluster_params
 * (opensfm/dense/depthmaps.py:154-175) sets from the config defaults
 * (opensfm/config.py:407-466): iterations 4, patch 5, sigma_spatial 1.5,
 * sigma_color 25/255, top_k 8 (depthmap_num_matching_views), census on,
 * 3 levels, smooth 0.2, anchors 2, filter off, speckle 0, gap 0; the other
 * fields as pm_default_params.  (The pipeline's geometric-consistency
 * weight, 0.05, goes through pm_set_geom_consistency.) */
void pm_pipeline_params(pm_params* out);

/* Last error message of the calling thread ("" if none). */
const char* pm_last_error(void);

int pm_device_count(int* count);
int pm_device_name(int device, char* buf, int buflen);
/* Total and currently free device memory in bytes (GlobalMemSize of
 * DepthmapClusterEstimator.device_memory_bytes, pybind.cc:105-122).  Either
 * pointer may be NULL. */
int pm_device_memory(int device, uint64_t* total_bytes, uint64_t* free_bytes);

int pm_create(int device, pm_handle** out);
int pm_destroy(pm_handle* h);

/* K, R: row-major 3x3 float64; t: 3 float64.  The image is copied.
 * row_stride in bytes (0 = packed).  mask may be NULL; it is accepted for
 * the pydense surface and not used (the reference ACMMP has no mask). */
int pm_add_view(pm_handle* h, const double* K, const double* R, const double* t,
                const void* image, int format, int width, int height, int64_t row_stride,
                const uint8_t* mask);
/* pm_add_view for an image already in device memory of the handle's device
 * (row_stride in bytes, 0 = packed): AddView (estimator.cc:33-42) for
 * callers whose views are resident in HBM.  The image is not copied: it is
 * read at the next pm_upload (or batch staging), which pads and converts it
 * on the device exactly as the host path does, so it must stay valid and
 * unchanged until then.  Such views run at one pyramid level. */
int pm_add_view_device(pm_handle* h, const double* K, const double* R, const double* t,
                       const void* device_image, int format, int width, int height, int64_t row_stride);
int pm_clear_views(pm_handle* h);
int pm_num_views(pm_handle* h, int* n);

int pm_set_depth_range(pm_handle* h, float depth_min, float depth_max);
int pm_set_params(pm_handle* h, const pm_params* p);
int pm_get_params(pm_handle* h, pm_params* p);

/* Geometric consistency: weight and, per source slot, a previous depth map
 * at the reference resolution (NULL entry = none); prev_depths = NULL sets
 * the weight only.  Coarser pyramid levels use the maps bilinearly resized
 * (UploadPreviousDepths, estimator.cc:66-116). */
int pm_set_geom_consistency(pm_handle* h, float weight, const float* const* prev_depths, int n);

/* One previous depth map for source slot `slot` (0-based), any size; it is
 * resized to each level bilinearly.  depth = NULL clears the slot.
 * SetPreviousDepths / ClearPreviousDepths (estimator.cc:59-64); the weight
 * comes from pm_set_geom_consistency (prev_depths may be NULL there). */
int pm_set_previous_depth(pm_handle* h, int slot, const float* depth, int width, int height);
int pm_clear_previous_depths(pm_handle* h);

/* Planar prior: (nx, ny, nz, d) per reference pixel + mask (NULL clears).
 * Planes live in the camera frame, so a coarser pyramid level takes the
 * prior of the full-resolution pixel at the centre of its footprint. */
int pm_set_prior(pm_handle* h, const float* prior_planes, const uint32_t* masks);

/* Stage views/cameras in device memory and allocate the per-pixel state
 * (full resolution; with several hierarchy levels pm_run stages each level
 * itself and ends with the full-resolution level staged). */
int pm_upload(pm_handle* h);
/* Per hierarchy level, coarse to fine: random init (coarsest) or transfer
 * from the decoded coarser level, prior re-init, iterations, optional
 * filter and decode, all on the device; returns once the stream is idle. */
int pm_run(pm_handle* h);
/* One pyramid level of params.hierarchy_levels (RunLevel, estimator.cc:
 * 172-255), for callers that feed other views' intermediate depths between
 * levels (DepthmapClusterEstimator::Run, cluster_estimator.cc:82-156).
 * Levels run in order from 0; the decoded maps of the level can be read with
 * pm_download (their size from pm_level_size); the last level is full size. */
int pm_run_level(pm_handle* h, int level);
int pm_level_size(pm_handle* h, int* width, int* height);
/* Free the device memory of the handle except the decoded coarser level that
 * the next pm_run_level starts from (ReleaseGpuBuffers, estimator.cc:1420). */
int pm_release_device(pm_handle* h);

/* Copy the decoded maps out.  Any pointer may be NULL.
 * depth, cost, confidence: H*W float; plane: H*W*4 float (n, d);
 * normal: H*W*3 float; views: H*W uint32 selected-view bitmask. */
int pm_download(pm_handle* h, float* depth, float* plane, float* normal, float* cost,
                uint32_t* views, float* confidence);
/* pm_upload + pm_run + pm_download. */
int pm_compute(pm_handle* h, float* depth, float* plane, float* normal, float* cost,
               uint32_t* views, float* confidence);

/* Staged execution (after pm_upload). */
int pm_init_state(pm_handle* h);
/* Fronto-parallel plane sweep, then decode and post-processing, as pm_run
 * ends (after pm_upload).  For compute_brute_force() of the pydense stub
 * (pydense.pyi:53, whose num_planes comes from set_depth_range,
 * pydense.pyi:57); the reference tree has no implementation, so the method
 * is defined here: num_planes >= 1 depths log-spaced in
 * [depth_min, depth_max] (the range's geometric mean for one plane), normal
 * facing the camera, each scored like a random-init plane (stride-1 NCC,
 * fallbacks, top-k mean); per pixel the lowest cost wins, the first plane on
 * ties. */
int pm_plane_sweep(pm_handle* h, int num_planes);
int pm_iterate(pm_handle* h, int iteration, int color /* 0, 1 or -1 for both */);
int pm_filter(pm_handle* h);
int pm_decode(pm_handle* h);
int pm_read_state(pm_handle* h, float* planes, float* costs, uint32_t* rng, uint32_t* sel);
int pm_write_state(pm_handle* h, const float* planes, const float* costs, const uint32_t* rng,
                   const uint32_t* sel);

/* Score n (pixel, plane) pairs on the uploaded views: per-view costs (n x
 * num_sources, row-major), top-k initial cost (n) and view bitmask (n).
 * xy: 2n int32; planes: 4n float.  init_cost / sel may be NULL. */
int pm_score_planes(pm_handle* h, int n, const int32_t* xy, const float* planes, float* costs,
                    float* init_cost, uint32_t* sel);

/* pm_score_planes without any device allocation: the pairs and results are
 * staged in a fixed device array of the library (6 MiB per GPU, held until
 * the process ends) instead of buffers allocated per call.  No counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  n <= PM_NOALLOC_MAX_PAIRS (else
 * PM_E_INVALID).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (from several threads or handles) race.  Same arguments and
 * results as pm_score_planes. */
#define PM_NOALLOC_MAX_PAIRS 65536
int pm_score_planes_noalloc(pm_handle* h, int n, const int32_t* xy, const float* planes, float* costs,
                            float* init_cost, uint32_t* sel);

/* Differentiable photometric cost of per-pixel planes (no counterpart in the
 * reference; the refinement after PatchMatch).  Per reference pixel: depth
 * d, normal slopes (a, b) of the camera-facing normal (a, b, -1), and a
 * view mask; the cost is the mean over the masked views valid at the
 * pixel of 1 - NCC (the stride-1 bilateral NCC of compute_cost_vector,
 * opencl_kernels.h:471-499, on the same local affine warp, :401-461, and no
 * textureless fallback); grad = d cost / d(d, a, b) (3 per pixel, NULL =
 * cost only).  depth, cost: H*W; ab: 2*H*W; views: H*W; after pm_upload.
 * device_pointers = 1: every pointer is device memory of the handle's
 * device and the call returns when the results are written. */
int pm_ncc_cost_grad(pm_handle* h, const float* depth, const float* ab, const uint32_t* views, float* cost,
                     float* grad, int device_pointers);
/* Gauss-Newton refinement of the PatchMatch state (after pm_run): `steps`
 * damped steps (Levenberg-Marquardt factor `damping`) on the residuals
 * 1 - NCC of each pixel's selected views, a step kept only where it lowers
 * the pixel's cost; then decode again (pm_download reads the result). */
int pm_gauss_newton(pm_handle* h, int steps, float damping);
/* Device time of the last pm_gauss_newton's steps in milliseconds. */
int pm_last_refine_timing(pm_handle* h, float* ms);

/* Device time of the last pm_run in milliseconds (HIP events on the
 * handle's stream): the whole run, and the iterations (all colour passes). */
int pm_last_timing(pm_handle* h, float* run_ms, float* sweep_ms, int* sweep_launches);

/* Kernel stages of a run; per-stage device time of the last pm_run (sum
 * over its launches, HIP events bracketing each launch) and launch count. */
#define PM_STAGE_INIT 0         /* random init + initial costs (acmmp_random_init) */
#define PM_STAGE_CANDIDATES 1   /* cost vectors of 8 propagated candidates + current plane */
#define PM_STAGE_FALLBACK 2     /* textureless strided-NCC / census fallbacks */
#define PM_STAGE_SELECT 3       /* view weights, MC costs, candidate acceptance */
#define PM_STAGE_REFINE_COST 4  /* cost vectors of the refinement hypotheses */
#define PM_STAGE_REFINE 5       /* MC costs and acceptance of the refinements */
#define PM_STAGES 6
int pm_last_stage_timing(pm_handle* h, int stage, float* ms, int* launches);

/* Colour-tile widths of the kernels (tiles are width x 64/width pixels of
 * one colour): cost_tile_x for the cost kernels and the fallback (4, 8 or
 * 16), pixel_tile_x for the per-pixel kernels (16 or 32); 0 = the default
 * (4; 32).  Results do not depend on
 * them.  pm_get_tiles returns the widths in effect.  No counterpart in the
 * reference (its work-group shape is fixed by the OpenCL launch,
 * estimator.cc:619-680). */
int pm_set_tiles(pm_handle* h, int cost_tile_x, int pixel_tile_x);
int pm_get_tiles(pm_handle* h, int* cost_tile_x, int* pixel_tile_x);
/* At most waves_per_simd waves of the textureless-fallback kernel per SIMD
 * (1..8; 0 or -1 = no cap, the default).  Results do
 * not depend on it; a batch takes its first reference's setting.  No
 * counterpart in the reference. */
int pm_set_fallback_occupancy(pm_handle* h, int waves_per_simd);

/* ---- Batched runs ----------------------------------------------------
 * A batch runs many references (each a pm_handle with its views, parameters
 * and depth range) through one launch per kernel of the schedule, captured
 * once into a HIP graph.  It replaces the sequential reference loop of
 * DepthmapClusterEstimator::Run (cluster_estimator.cc:117-151) at one
 * pyramid level; every reference's result equals pm_run on that handle
 * alone, and lands in the handle (pm_download).  References may differ in
 * size, source count and depth range; patch size, image storage, iterations
 * and the checkerboard filter must be the same for all, hierarchy_levels 1.
 * The batch does not own the handles: they must outlive its runs.  A
 * change to a handle (views, parameters, tiles) is noticed by the next
 * pm_batch_run, which stages and captures again. */
typedef struct pm_batch pm_batch;
int pm_batch_create(int device, pm_batch** out);
int pm_batch_destroy(pm_batch* b);
int pm_batch_add(pm_batch* b, pm_handle* ref);
int pm_batch_clear(pm_batch* b);
int pm_batch_size(pm_batch* b, int* n);
/* cost_tile_x / pixel_tile_x: 0 = chosen by the autotuner; use_graph: 0 =
 * enqueue the launches on every run (per-stage timing available). */
int pm_batch_set_config(pm_batch* b, int cost_tile_x, int pixel_tile_x, int use_graph);
/* Pyramid level the batch runs (default 0): every reference is staged at
 * that level of its hierarchy_levels pyramid; level 0 starts from random
 * planes, finer levels from each reference's previous, coarser result, which
 * a run keeps for the next level.  The batched form of running RunLevel
 * level by level for every reference (estimator.cc:172-255,
 * cluster_estimator.cc:117-151); levels run in order. */
int pm_batch_set_level(pm_batch* b, int level);
/* Configuration in use after pm_batch_prepare / pm_batch_run, and the time
 * the autotuner took (0 when its result came from the per-process cache). */
int pm_batch_get_config(pm_batch* b, int* cost_tile_x, int* pixel_tile_x, float* tune_ms);
/* The autotuner's trials, 8 floats each: cost_tile_x, pixel_tile_x, then
 * the device ms of candidates, fallback, select, refine_cost, fallback,
 * refine.  *count = number of floats. */
int pm_batch_tune_log(pm_batch* b, float* out, int capacity, int* count);
/* Forget the autotuner's per-process results. */
void pm_batch_reset_tuning(void);
/* Stage all references, tune, capture the graph (pm_batch_run does this
 * when needed). */
int pm_batch_prepare(pm_batch* b);
int pm_batch_run(pm_batch* b);
/* Device time of the last pm_batch_run (ms). */
int pm_batch_timing(pm_batch* b, float* run_ms);
/* Per-stage device time of the last pm_batch_run without a graph
 * (use_graph 0; HIP events around every launch); zero after a graph run. */
int pm_batch_stage_timing(pm_batch* b, int stage, float* ms, int* launches);
/* Device memory (bytes) that staging and running this handle at its finest
 * level allocates; for admitting references to a batch within a memory
 * budget (the reference frees each reference's buffers after its run,
 * cluster_estimator.cc:150). */
int pm_device_bytes(pm_handle* h, uint64_t* bytes);
/* Keep up to max_bytes of the device buffers that pm_release_device and
 * pm_destroy give back, for reuse by later handles that need blocks of the
 * same sizes (0: free them at once, the default; lowering the limit frees
 * blocks).  No counterpart in the reference (which frees each reference's
 * buffers, cluster_estimator.cc:150); for runs of many references of one
 * shape in memory-admitted groups, where freeing and re-allocating the
 * buffers between groups costs seconds.  Cached blocks count as used device
 * memory; a failed allocation frees them all and tries again. */
int pm_set_buffer_cache(int device, uint64_t max_bytes);
/* Bytes held by the buffer cache of a device. */
int pm_buffer_cache_bytes(int device, uint64_t* held);

/* ---- SfM planar prior and neighbour selection --------------------------
 * Host-side steps around the PatchMatch path. */

/* SfM points (num_points x 3 float64, world frame) of the reference:
 * SetSfMPoints (estimator.cc:118-132).  Each pyramid level then builds its
 * planar prior from them (ComputeSfMPlanarPrior, estimator.cc:838-1107) and
 * PatchMatch re-seeds the masked pixels from it, as with pm_set_prior
 * (which this replaces; points = NULL / num_points = 0 clears them). */
int pm_set_sfm_points(pm_handle* h, const double* points, int num_points);
/* ComputeSfMPlanarPrior for a w x h image with intrinsics K (float64
 * row-major, of that image size), R, t: prior_planes (4 floats per pixel,
 * (0, 0, -1, 0) where unmasked) and masks (1 where a triangle's plane gives
 * a depth in [depth_min, depth_max]); *num_triangles = triangles kept after
 * the depth-ratio (1.5) and area (20000 px) gates.  The triangulation is a
 * Delaunay triangulation in canonical order (cv::Subdiv2D in the
 * reference). */
int pm_sfm_planar_prior(const double* K, const double* R, const double* t, int width, int height, float depth_min,
                        float depth_max, const double* points, int num_points, float* prior_planes, uint32_t* masks,
                        int* num_triangles);
/* Delaunay triangulation of n 2-D points (xy interleaved), canonical order
 * (vertex indices ascending, triangles sorted); writes min(capacity,
 * count) triangles and sets *count.  No counterpart in the reference (it
 * calls cv::Subdiv2D); exposed for testing the prior. */
int pm_delaunay(const double* xy, int n, int* triangles, int capacity, int* count);
/* SelectNeighbors (opensfm/src/lib/sfm/src/dense_helpers.cc:331-568):
 * pair scores of num_shots cameras (origins 3 x float64 each) over
 * num_points super-points (centroid, track count, the shots that see it:
 * vis_shots[vis_offsets[p] .. vis_offsets[p + 1])).  best: num_shots x
 * (1 + num_neighbors) shot indices (self first, then by score), all_out:
 * num_shots x num_shots (best, then every other shot sharing more than
 * min_point_all tracks), both -1 padded, with counts. */
int pm_select_neighbors(int num_shots, const double* origins, int num_points, const double* coords,
                        const int* vis_offsets, const int* vis_shots, const int* track_counts, int num_neighbors,
                        int min_point_best, int min_point_all, double theta_min_deg, double theta_max_deg, int* best,
                        int* best_counts, int* all_out, int* all_counts);

/* ---- Depth-map cleaner ----------------------------------------------
 * Cross-view depth consistency with space carving and a grazing-angle rule:
 * GPUDepthmapCleaner (opensfm/src/lib/dense/cleaner.h:21-102, cleaner.cc)
 * and its kernel acmmp_clean_depthmap (opencl_kernels.h:2116-2279), bound
 * to Python as GPUDepthmapCleaner (python/pybind.cc:125-143).  Depth maps
 * stay in device memory from pm_cleaner_add_view until pm_cleaner_clear. */
typedef struct pm_cleaner pm_cleaner;
int pm_cleaner_create(int device, pm_cleaner** out);
int pm_cleaner_destroy(pm_cleaner* c);
/* SetSameDepthThreshold, SetMinConsistentViews, SetCarvingThreshold,
 * SetMaxCarvedViews, SetGrazingCosThreshold (cleaner.h:27-43; defaults
 * 0.01, 3, 0.05, 2, 0.2). */
int pm_cleaner_set_params(pm_cleaner* c, float same_depth_threshold, int min_consistent_views,
                          float carving_threshold, int max_carved_views, float grazing_cos_threshold);
/* AddView (cleaner.h:46-51): K, R row-major float64, t float64, depth
 * (height x width float32, copied); normal: camera-frame normals, 3 floats
 * per pixel interleaved, or NULL.  *index = the view's index. */
int pm_cleaner_add_view(pm_cleaner* c, const double* K, const double* R, const double* t, const float* depth,
                        int width, int height, const float* normal, int* index);
int pm_cleaner_num_views(pm_cleaner* c, int* n);
int pm_cleaner_view_size(pm_cleaner* c, int index, int* width, int* height);
/* Clean (cleaner.cc:95-242): the cleaned depth of view ref_idx (its size)
 * against up to 16 neighbours, into out (caller-owned, width x height). */
int pm_cleaner_clean(pm_cleaner* c, int ref_idx, const int* neighbor_ids, int num_neighbors, float* out);
/* Device time of the last pm_cleaner_clean's kernel (ms). */
int pm_cleaner_last_timing(pm_cleaner* c, float* ms);
/* Clear (cleaner.cc:244-256): drop every view and its device memory. */
int pm_cleaner_clear(pm_cleaner* c);

#ifdef __cplusplus
}
#endif

#endif /* OPENSFM_PM_H_ */
