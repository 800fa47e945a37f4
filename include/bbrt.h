/* bbrt.h -- C ABI of libbbrt_hip.so, the MI355X (gfx950) billiard bounce loop.
 *
 * The library replaces the reference's per-ray trace loop,
 *   RayTracingDriver::trace_ray        (include/RayTracingDriver.hpp:172-242)
 * as driven by
 *   RayTracingDriver::run              (include/RayTracingDriver.hpp:143-150),
 *   PlaneSourceDriver::run             (include/PlaneSourceDriver.hpp:79-109),
 *   BounceMapDriver::run               (include/BounceMapDriver.hpp:202-238),
 * including everything below it: BVH::intersect (include/BVH.hpp:75-128),
 * SurfacePatch::intersect (include/ScatObjects/SurfacePatch.hpp:253-303),
 * moller_trumbore (include/Utils/MollerTrumbore.hpp:32-64) and newton_refine
 * (SurfacePatch.hpp:593-660).  Scene construction (SurfacePatch's constructor,
 * SurfacePatch.hpp:125-131, and RayTracingDriver::build_bvh,
 * RayTracingDriver.hpp:375-384) happens in bbrt_scene_create.
 *
 * Conventions: every call returns 0 on success or a negative code, with a
 * message in bbrt_last_error() (thread-local).  No C++ exception crosses the
 * ABI.  A scene handle owns all of its device memory and one HIP stream; a
 * handle is not reentrant, distinct handles may be used from distinct threads.
 */
#ifndef BBRT_H
#define BBRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bbrt_scene_s* bbrt_scene;

/* Tensor-product patches in patch-index order.  The reference's patch index is
 * the lexicographic order of the patch file names (RayTracingDriver.hpp:361,
 * :379-382); callers that load files pass them in that order.
 *   imax, jmax   [n]     node counts per patch (SurfacePatch::Data, SurfacePatch.hpp:60-61)
 *   u_off, v_off [n+1]   offsets of patch k's u and v nodes in unodes / vnodes
 *   f_off        [n+1]   node offsets: patch k's 12 fields start at fields + 12*f_off[k],
 *                        laid out [field][i*jmax + j] in the order
 *                        x y z dxdu dydu dzdu dxdv dydv dzdv nuX nuY nuZ
 *                        (SurfacePatch.hpp:74-80; dS and mask are not read). */
typedef struct {
    int64_t n_patches;
    const int32_t* imax;
    const int32_t* jmax;
    const int64_t* u_off;
    const int64_t* v_off;
    const int64_t* f_off;
    const double* unodes;
    const double* vnodes;
    const double* fields;
} bbrt_patch_set;

enum { BBRT_NEWTON_REF = 0,          /* the reference stopping rule (step < 1e-3), SurfacePatch.hpp:598-630 */
       BBRT_NEWTON_CONVERGED = 1,    /* same classification, then FP64 polish to step < 1e-13*max(1,|t|) */
       BBRT_NEWTON_CONVERGED_FP32 = 2 };  /* BBRT_NEWTON_CONVERGED whose first iterations run in FP32: FP32 from
                                           the seed until its step is < 1e-3, then the FP64 iterations (the
                                           reference's stop, determinant and domain decisions, then the polish);
                                           when the FP32 phase fails, FP64 from the seed.  Classifications can
                                           differ from BBRT_NEWTON_CONVERGED where FP32 and FP64 iterations part
                                           ways (measured, not promised).  Traces in the persistent schedule
                                           only, without ray_work */
enum { BBRT_SEED_TRIANGLE = 0, BBRT_SEED_LEARNED = 1 };
/* RayPath::Status, RayTracingDriver.hpp:69 */
enum { BBRT_ESCAPED = 0, BBRT_TRAPPED = 1, BBRT_ERROR = 2 };
/* Error kinds: the throw sites SurfacePatch.hpp:618-622 (singular), :632-637
 * (no convergence), :639-646 (domain), RayTracingDriver.hpp:220-226 (back
 * face), :176-181 (zero direction); STACK = traversal stack overflow. */
enum { BBRT_ERR_NONE = 0, BBRT_ERR_SINGULAR = 1, BBRT_ERR_NOCONV = 2, BBRT_ERR_DOMAIN = 3,
       BBRT_ERR_BACKFACE = 4, BBRT_ERR_ZERO_DIR = 5, BBRT_ERR_STACK = 6 };
enum { BBRT_FLAG_NO_CULL = 1u,          /* stage 1 as the reference's brute force (all 58,806 tests) */
       BBRT_FLAG_COUNT_DIRECT_HIT = 2u, /* hit_count also counts each ray's first hit (PlaneSourceDriver.hpp:101) */
       BBRT_FLAG_TIME_KERNELS = 4u,     /* record HIP events around every bounce-kernel launch */
       BBRT_FLAG_WAVEFRONT = 8u,        /* one launch per bounce with live-ray compaction instead of the
                                           persistent kernel with a device ray queue */
       BBRT_FLAG_GRAPH = 16u,           /* capture the trace (counter fills, k_init, persistent kernel) in a
                                           hipGraph on the handle's stream and replay it while the arguments
                                           are unchanged; kernel_ms[0] is then the whole graph's time */
       BBRT_FLAG_STAGE1_WAVE = 32u,     /* persistent kernel, stored scenes: every stage 1 walked by the whole
                                           wave, one lane's (ray, patch) pair at a time */
       BBRT_FLAG_STAGE1_LANE = 64u,     /* ... every stage 1 walked by its own lane (default: the wave walks all
                                           its leaf lanes' pyramids together, breadth-first, one box group or
                                           cell a lane per pass); results and counters do not change.
                                           Compact scenes ignore both */
       BBRT_FLAG_SEED_FIRST = 256u,     /* learned seed, network first (opt-in, parity measured, not promised):
                                           each candidate patch goes to Newton from the network's seed without
                                           stage 1.  A root that converges inside the patch (classification and
                                           polish) is the hit; one that converges outside the domain by more
                                           than BBRT_SEED_FIRST_MISS_MARGIN in u or v is a miss; anything else
                                           runs stage 1 and the triangle seed, whose outcome stands.  Stage 1
                                           is then skipped for most candidates, but a patch the reference's
                                           stage 1 misses can be hit and a farther root can be taken, so the
                                           records can differ from the reference's */
       BBRT_FLAG_GROUP_RAYS = 128u };   /* persistent kernel: take the live rays grouped by where they pass
                                           their scene (Morton order of the cells of the root box: 32^3 for a
                                           one-scene handle, 4^3 per sub-scene of a batch handle, sub-scene
                                           major; above 2^22 bins, i.e. more than 65,536 sub-scenes, the flag
                                           is ignored and the queue keeps input order) instead of in input
                                           order.  Each wave then traces rays that
                                           pass the same part of the scene.  Pays with several traces in flight
                                           (bbrt_trace_device_async); one trace at a time it can be slower,
                                           since the expensive rays then come together.  Results and counters
                                           do not change */

#define BBRT_SEED_FIRST_MISS_MARGIN 0.25

typedef struct {
    int32_t newton_mode;   /* BBRT_NEWTON_* */
    int32_t seed_mode;     /* BBRT_SEED_*; LEARNED needs BBRT_NEWTON_CONVERGED and bbrt_scene_set_mlp */
    int32_t max_bounces;   /* run(max_bounces), RayTracingDriver.hpp:143 */
    int32_t record_k;      /* bounces recorded per ray in hit_patch / hit_xyz / newton_iters */
    uint32_t flags;        /* BBRT_FLAG_* */
    int32_t block;         /* threads per block of the bounce kernel; 0 = library default (64 for the
                              persistent kernel on a stored scene with the triangle seed, else 256) */
    int32_t waves_per_eu;  /* occupancy target of the persistent kernel: 0 = library default (2), or 2, 3, 4
                              (the compiler keeps registers within that occupancy, spilling if needed) */
    float grid_fraction;   /* persistent kernel: fraction of the resident workgroups launched (0 = all) */
    int32_t newton_batch;  /* persistent kernel: lanes of a wave whose stage 1 hit wait until this many can run
                              Newton (and the learned seed's network) together (1..64; 0 = library default:
                              64, or 32 on a compact scene); results do not depend on it */
} bbrt_params;

/* Outputs; any pointer may be NULL.  Host memory for bbrt_trace, device
 * memory for bbrt_trace_device.
 *   status, err_kind       [n] int8    RayPath::status and the error kind
 *   bounces                [n] int32   RayPath::bounce_count
 *   err_patch              [n] int32   patch whose intersection failed, -1 otherwise
 *   hit_patch              [n][record_k] int32  RayPath::hit_objects as patch indices, -1 padded
 *   hit_xyz                [n][record_k][3] double  RayPath::points[1..]
 *   newton_iters           [n][record_k] uint8  Newton iterations of each recorded hit
 *   final_od               [n][6] double  origin and unit direction when the ray stopped
 *                          (an escaped ray's tail point is origin + dir * scene diagonal,
 *                          RayTracingDriver.hpp:198-203)
 *   hit_count              [n_patches] int64  PlaneSourceDriver hit counts (accumulated,
 *                          not cleared), first hit skipped unless BBRT_FLAG_COUNT_DIRECT_HIT
 * Record slots past a ray's bounce count are written only by bbrt_trace (which
 * pads them with -1 / 0); bbrt_trace_device leaves them untouched, like the
 * reference's RayPath vectors, which hold bounce_count entries. */
typedef struct {
    int8_t* status;
    int8_t* err_kind;
    int32_t* bounces;
    int32_t* err_patch;
    int32_t* hit_patch;
    double* hit_xyz;
    uint8_t* newton_iters;
    double* final_od;
    int64_t* hit_count;
    uint32_t* ray_work;   /* [n][4]: BVH node tests, culling box tests, Moller-Trumbore tests and
                             Newton iterations spent on each ray, summed over its bounces */
    uint32_t* seed_stats; /* [n][2], learned-seed mode: the ray's stage-1 hits whose learned-seed root was
                             kept, and those that fell back to the triangle seed (zeros otherwise) */
} bbrt_outputs;

/* Work counters returned by the trace calls (int64[BBRT_NCOUNTERS]). */
enum { BBRT_CTR_RAYS = 0, BBRT_CTR_BOUNCES = 1, BBRT_CTR_BVH_NODE_TESTS = 2, BBRT_CTR_CANDIDATES = 3,
       BBRT_CTR_CULL_BOX_TESTS = 4, BBRT_CTR_MT_TESTS = 5, BBRT_CTR_NEWTON_CALLS = 6, BBRT_CTR_NEWTON_ITERS = 7,
       BBRT_CTR_NEWTON_FLOPS = 8, BBRT_CTR_POLISH_FAIL = 9, BBRT_CTR_LAUNCHES = 10,
       /* learned-seed mode: stage-1 hits whose learned-seed Newton was accepted, those that fell back to the
          triangle seed, wave-level network evaluations, and lanes seeded by them */
       BBRT_CTR_SEED_LEARNED = 11, BBRT_CTR_SEED_FALLBACK = 12, BBRT_CTR_MLP_EVALS = 13, BBRT_CTR_MLP_LANES = 14,
       /* BBRT_NEWTON_CONVERGED_FP32: FP32 Newton iterations (not in NEWTON_ITERS or NEWTON_FLOPS, which
          count the FP64 work) */
       BBRT_CTR_NEWTON_ITERS_FP32 = 15,
       /* BBRT_FLAG_SEED_FIRST: candidates the network seed's Newton declared missed (converged outside the
          patch by more than BBRT_SEED_FIRST_MISS_MARGIN), without stage 1 */
       BBRT_CTR_SEED_MISS = 16,
       BBRT_NCOUNTERS = 20 };

/* Builds a scene: per patch the barycentric weights, centroid, node box,
 * 100x100 coarse grid, per-cell sagitta and normal and the inflated box
 * (SurfacePatch.hpp:125-131, :335-336, :380-492); then the median-split BVH
 * with leaf threshold 4 (BVH.hpp:136-189) and the scene diagonal
 * (RayTracingDriver.hpp:386-396); then uploads everything to the device. */
int bbrt_scene_create(const bbrt_patch_set* patches, bbrt_scene* out);
void bbrt_scene_destroy(bbrt_scene scene);

/* Scene scalars: scene diagonal, largest patch sagitta, BVH node count, patch
 * count and device bytes held. Any pointer may be NULL. */
int bbrt_scene_info(bbrt_scene scene, double* diag, double* max_patch_sagitta, int64_t* n_nodes,
                    int64_t* n_patches, int64_t* device_bytes);
/* Per patch: inflated box [n][6] (min xyz, max xyz), centroid [n][3], max sagitta [n]. */
int bbrt_scene_patch_info(bbrt_scene scene, double* box, double* centroid, double* max_sagitta);
/* The flattened BVH in the reference's depth-first layout (BVH.hpp:56-62):
 * node boxes [m][6], right child (-1 = leaf), leaf object start and count,
 * split axis, and the leaf object array [n_patches]. */
int bbrt_scene_bvh(bbrt_scene scene, double* node_box, int32_t* right, int32_t* start, int32_t* count,
                   uint8_t* axis, int32_t* leaf_objects);
/* Coarse grid of patch k: vertices [100*100][3], cell sagitta [99*99], cell normals [99*99][3]. */
int bbrt_scene_coarse(bbrt_scene scene, int64_t k, double* xyz, double* sagitta, double* cell_normal);

/* Traces n rays (host memory in and out) -- RayTracingDriver::set_rays + run
 * (RayTracingDriver.hpp:119-150) for the whole batch.  counters may be NULL. */
int bbrt_trace(bbrt_scene scene, const double* origins, const double* dirs, int64_t n, const bbrt_params* params,
               const bbrt_outputs* out, int64_t* counters);

/* Same with origins, dirs and outputs in device memory, on `stream` (a
 * hipStream_t; NULL = the scene's own stream).  Returns after the work is
 * done.  kernel_ms (may be NULL) receives [0] the summed duration of the
 * bounce-kernel launches (BBRT_FLAG_TIME_KERNELS), [1] the launch count,
 * [2] the wall time of the whole call on the stream, in ms. */
int bbrt_trace_device(bbrt_scene scene, const double* origins, const double* dirs, int64_t n,
                      const bbrt_params* params, const bbrt_outputs* out, int64_t* counters, void* stream,
                      double* kernel_ms);

/* Several independent scenes in one handle (the batched bounce map): sub-scene
 * i is built from sets[i] exactly as bbrt_scene_create builds a scene (own
 * BVH, diagonal and culling limit); its patches get global indices
 * patch_begin[i] .. patch_begin[i+1]-1.  No counterpart in the reference,
 * whose drivers trace one scene at a time (RayTracingDriver.hpp:119-150). */
int bbrt_scene_create_batch(const bbrt_patch_set* sets, int32_t n_sets, bbrt_scene* out);
/* bbrt_scene_create_batch with scene flags.  BBRT_SCENE_COMPACT keeps neither
 * the 100x100 coarse grid, the cell sagittas and normals nor the FP32 cell
 * boxes (~800 KB per patch): the kernels evaluate each coarse cell a ray
 * reaches from the nodal fields and small per-node-set tables, bit-identical
 * to the stored values, so records and counters are the same.  About 33 KB per
 * order-4 patch stays resident.  bbrt_scene_coarse is unavailable on it. */
enum { BBRT_SCENE_COMPACT = 1u,
       BBRT_SCENE_DEVICE_BUILD = 2u };   /* with COMPACT: the per-cell work of the build (coarse grid, cell
                                           padding, culling boxes) and every sub-scene's BVH (with its 4-wide
                                           traversal nodes) are built on the GPU, bit-identical to the host
                                           build; the host keeps patch records, centroids and node boxes */
int bbrt_scene_create_ex(const bbrt_patch_set* sets, int32_t n_sets, uint32_t scene_flags, bbrt_scene* out);
/* Sub-scene count, patch_begin [n+1], BVH root node [n], diagonal [n]; any pointer may be NULL. */
int bbrt_scene_batch_info(bbrt_scene scene, int32_t* n_scenes, int64_t* patch_begin, int32_t* bvh_root,
                          double* diag);
/* Traces all sub-scenes' rays in one persistent launch.  ray_begin [n_scenes+1]
 * (host array): rays ray_begin[i] .. ray_begin[i+1]-1 belong to sub-scene i,
 * ray_begin[0] = 0.  Outputs are per ray as for bbrt_trace; hit_patch and
 * err_patch hold sub-scene-local patch indices, hit_count is indexed by
 * global patch index.  bbrt_trace / bbrt_trace_device refuse a handle with
 * more than one sub-scene. */
int bbrt_trace_batch(bbrt_scene scene, const double* origins, const double* dirs, const int64_t* ray_begin,
                     const bbrt_params* params, const bbrt_outputs* out, int64_t* counters);
int bbrt_trace_batch_device(bbrt_scene scene, const double* origins, const double* dirs, const int64_t* ray_begin,
                            const bbrt_params* params, const bbrt_outputs* out, int64_t* counters, void* stream,
                            double* kernel_ms);

/* bbrt_trace_device / bbrt_trace_batch_device without waiting: enqueues the
 * trace on `stream` (NULL = the scene's own stream) and returns.  Replaces the
 * same ray loops as bbrt_trace_device (RayTracingDriver.hpp:119-150); the
 * reference has no asynchronous form.
 * ray_begin: NULL for a one-scene handle, else as for bbrt_trace_batch (host
 * array, read during the call), with n = ray_begin[n_scenes].
 * d_counters: NULL or device memory for BBRT_NCOUNTERS int64, written on the
 * stream when the trace ends.
 * Each stream gets its own scratch in the handle, so traces enqueued on
 * distinct streams may run at the same time: the next batch's rays start on
 * CUs the previous batch's last rays have left.  Calls on one stream run in
 * order.  Inputs and outputs must stay valid, and the handle alive, until the
 * stream has finished the trace.  Calls may come from several host threads
 * (they serialise on the handle for the enqueue only), but not concurrently
 * with the handle's synchronous calls.  BBRT_FLAG_WAVEFRONT, BBRT_FLAG_GRAPH
 * and BBRT_FLAG_TIME_KERNELS are refused (they need the host to wait).
 * Changing ray_begin between calls waits for the whole device once. */
int bbrt_trace_device_async(bbrt_scene scene, const double* origins, const double* dirs, int64_t n,
                            const int64_t* ray_begin, const bbrt_params* params, const bbrt_outputs* out,
                            int64_t* d_counters, void* stream);

/* bbrt_trace_device_async that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The trace's scratch (ray state, live lists, queue bins, counts, counters: 56
 * bytes per ray, 60 with BBRT_FLAG_GROUP_RAYS, plus up to 256 KB) lives in one
 * file-scope device array of BBRT_NOALLOC_SCRATCH_BYTES per GPU, which each GPU
 * the library runs on holds until the process ends; traces needing more are
 * refused with an error (2^20 rays fit in every mode).  The call clears what
 * must start at zero, enqueues the trace on `stream` (NULL = the scene's own
 * stream) and returns; counters go to d_counters (device memory, may be NULL).
 * Every call on one GPU shares that GPU's array: calls whose traces overlap on
 * one GPU race (enqueue them on one stream, or wait between them).  One-scene
 * handles only; BBRT_FLAG_WAVEFRONT, BBRT_FLAG_GRAPH and BBRT_FLAG_TIME_KERNELS
 * are refused, as for bbrt_trace_device_async.  Results and counters equal
 * bbrt_trace_device's. */
#define BBRT_NOALLOC_SCRATCH_BYTES (64u << 20)
int bbrt_trace_device_noalloc(bbrt_scene scene, const double* origins, const double* dirs, int64_t n,
                              const bbrt_params* params, const bbrt_outputs* out, int64_t* d_counters, void* stream);

/* Launch-shape autotuner for the persistent kernel.  Times the trace of the
 * given device-resident rays (ray_begin as for bbrt_trace_batch, or NULL and n
 * rays of a single scene; outputs in `out` are overwritten) under a coordinate
 * descent from the library defaults of the kernel the trace selects (block 64
 * for a stored scene with the triangle seed, else 256; 2 waves/EU; Newton
 * batch 64, 32 on a compact scene; the full grid): waves_per_eu {2, 3, 4}, newton_batch
 * {8, 16, 32, 64}, block {64, 128, 256}, grid_fraction {0.5, 0.75, 1}; each
 * candidate is the faster of two runs.  *best = *base with the fastest shape;
 * table (may be NULL) receives up to max_entries timed candidates in order. */
typedef struct {
    int32_t block, waves_per_eu, newton_batch;
    float grid_fraction;
    double ms;
} bbrt_tune_entry;
int bbrt_autotune(bbrt_scene scene, const double* origins, const double* dirs, const int64_t* ray_begin, int64_t n,
                  const bbrt_params* base, const bbrt_outputs* out, bbrt_params* best, bbrt_tune_entry* table,
                  int32_t max_entries, int32_t* n_entries, void* stream);

/* Host-only scene preparation (no device needed): what bbrt_scene_create
 * computes per patch.  box [n][6], centroid [n][3], max_sagitta [n],
 * coarse_xyz [n][100*100][3], cell [n][99*99][4] (sagitta, normal xyz); any
 * output may be NULL.  Mirrors SurfacePatch's constructor
 * (SurfacePatch.hpp:125-131, :380-492). */
int bbrt_prepare_patches(const bbrt_patch_set* patches, double* box, double* centroid, double* max_sagitta,
                         double* coarse_xyz, double* cell);

/* Host-only BVH build over arbitrary objects, BVH::BVH(objects,
 * leaf_threshold) (BVH.hpp:67-73, :136-189): boxes [n][6], centroids [n][3].
 * Outputs are sized [2n] (nodes) and [n] (leaf objects); returns the node
 * count, or -1 on error. */
int64_t bbrt_build_bvh(const double* box, const double* centroid, int64_t n, int32_t leaf_threshold,
                       double* node_box, int32_t* right, int32_t* start, int32_t* count, uint8_t* axis,
                       int32_t* leaf_objects);

/* bbrt_build_bvh with the build on the GPU (host arrays in and out): the same
 * tree, bit for bit -- nodes, boxes, split axes, child links and the leaf
 * order that libstdc++'s std::nth_element leaves (BVH.hpp:166-179).  The
 * device build of scene handles (BBRT_SCENE_DEVICE_BUILD) runs the same code. */
int64_t bbrt_build_bvh_gpu(const double* box, const double* centroid, int64_t n, int32_t leaf_threshold,
                           double* node_box, int32_t* right, int32_t* start, int32_t* count, uint8_t* axis,
                           int32_t* leaf_objects);

/* BounceMapDriver::run (include/BounceMapDriver.hpp:103-259) on the device for
 * n_sample patches of a one-scene handle (patches [n_sample], host array of
 * patch indices): per patch an n x n grid of surface points (sample_grid,
 * SurfacePatch.hpp:199-242, u, v = -1 + 2 i / (n - 1)), each launched from its
 * point nudged along the normal by eps[patch] (the driver's
 * self_intersect_epsilon of that patch, BounceMapDriver.hpp:161) into a
 * theta_n x phi_n solid-angle-uniform cone about the normal in the tangent
 * frame of build_tangent_frame (:202-220, :499-519); rays point-major, theta
 * then phi.  cos_theta_max and the cos/sin tables of the phi_n angles
 * (j + 0.5) 2 pi / phi_n come from the caller's libm, so the launch rays equal
 * a host computation's bit for bit.  All rays are traced in one persistent
 * launch; per point the largest bounce count of the directions that did not
 * end in an Error, the Error count, and the first direction reaching that
 * count (:214-236) are reduced on the device. */
typedef struct {
    int32_t grid_n;        /* position_grid_n, >= 2 */
    int32_t theta_n, phi_n;
    int32_t max_bounces;
    int32_t newton_mode;   /* BBRT_NEWTON_* */
    uint32_t flags;        /* BBRT_FLAG_* of the trace (graph and kernel timing are ignored) */
    double cos_theta_max;  /* cos(cone_angle_deg / 2) */
    const double* cos_phi; /* [phi_n] */
    const double* sin_phi; /* [phi_n] */
    const double* eps;     /* [n_sample] launch nudge of each sampled patch */
} bbrt_bounce_map_params;
/* Host outputs, any may be NULL: grid [n_sample][n*n][9] (point, sample_grid's
 * unit normal, dX/du); max_bounces, error_count [n_sample][n*n]; best_dir
 * [n_sample][n*n] (theta * phi_n + phi, -1 when every direction errored);
 * best_od [n_sample][n*n][6] (that direction's launch origin and direction);
 * launch_o, launch_d [n_sample*n*n*theta_n*phi_n][3]. */
typedef struct {
    double* grid;
    int32_t* max_bounces;
    int32_t* error_count;
    int32_t* best_dir;
    double* best_od;
    double* launch_o;
    double* launch_d;
} bbrt_bounce_map_outputs;
int bbrt_bounce_map(bbrt_scene scene, const int64_t* patches, int64_t n_sample, const bbrt_bounce_map_params* params,
                    const bbrt_bounce_map_outputs* out, int64_t* counters);

/* ------------------------------------------------------------------------
 * Object-level intersection, for callers that intersect objects directly
 * instead of tracing rays through a driver (the reference's own test
 * tests/Triangle_Intersect.cpp does).  Host arrays in and out; the work runs on
 * the GPU with the bounce loop's arithmetic.
 * ------------------------------------------------------------------------ */
/* Moller-Trumbore on n (ray, triangle) pairs: moller_trumbore(ray, v0, v1, v2,
 * eps) (include/Utils/MollerTrumbore.hpp:32-64), as Triangle::intersect calls
 * it (include/ScatObjects/Triangle.hpp:29-31).  tri [n][9] = v0, v1, v2;
 * origins, dirs [n][3].  Out: hit [n] (1/0) and for hits t [n], point [n][3]
 * (o + t d) and the unit face normal [n][3]. */
int bbrt_moller_trumbore(const double* tri, const double* origins, const double* dirs, int64_t n, double eps,
                         int8_t* hit, double* t, double* point, double* normal);

/* A set of m flat triangles on the device with its BVH, built as
 * BVH::BVH(objects, leaf_threshold) builds it (include/BVH.hpp:67-73,
 * :136-189) from the triangles' boxes and centroids (Triangle.hpp:11-23).
 * tri [m][9]. */
typedef struct bbrt_triangles_s* bbrt_triangles;
int bbrt_triangles_create(const double* tri, int64_t m, int32_t leaf_threshold, bbrt_triangles* out);
void bbrt_triangles_destroy(bbrt_triangles triangles);
/* BVH::intersect (BVH.hpp:75-128) over the set for n rays, or, with
 * brute_force, brute_force_intersect (include/BruteForce.hpp:7-15) over the
 * triangles in order.  Out: object [n] (the triangle index, -1 for a miss) and
 * for hits t, point [n][3], normal [n][3]. */
int bbrt_triangles_intersect(bbrt_triangles triangles, const double* origins, const double* dirs, int64_t n,
                             int32_t brute_force, int32_t* object, double* t, double* point, double* normal);

/* BVH::intersect (BVH.hpp:75-128) over a one-scene handle's patches, with
 * SurfacePatch::intersect (include/ScatObjects/SurfacePatch.hpp:253-303) at the
 * leaves, for n rays whose directions are used as given; or, with brute_force,
 * brute_force_intersect (BruteForce.hpp:7-15) over the patches in index order.
 * Out per ray: patch (-1: no hit or an error), t, point [3], normal [3] of the
 * closest hit, and err_kind (BBRT_ERR_*) / err_patch of the intersection that
 * threw, which ends the ray's query as the exception ends BVH::intersect. */
int bbrt_scene_intersect(bbrt_scene scene, const double* origins, const double* dirs, int64_t n, int32_t newton_mode,
                         int32_t brute_force, int32_t* patch, double* t, double* point, double* normal,
                         int8_t* err_kind, int32_t* err_patch);

/* ------------------------------------------------------------------------
 * Learned Newton seed.  No counterpart in the reference, whose Newton seed is
 * always the centroid of the closest stage-1 triangle
 * (SurfacePatch.hpp:272-303, :499-510).  A 51 -> 64 -> 64 -> 64 -> 3 GELU
 * network predicts the converged (u, v, t) of a (ray, patch) pair from the ray
 * in the patch's local frame (6 values) and 15 points of the patch (45
 * values).  It is trained on the GPU (bf16 matrix cores, fp32 accumulation,
 * Adam on fp32 master weights) on pairs that the method itself solves.
 * With seed_mode BBRT_SEED_LEARNED (converged Newton mode only) stage 1 still
 * decides hit or miss; for a hit, Newton starts from the network's seed, and
 * its result is kept only if it converges (classification and polish) to a
 * root within two coarse cells of the triangle seed.  Otherwise Newton runs
 * again from the triangle seed, whose outcome stands -- so every result is
 * one the triangle-seed path can produce, whatever the network says.
 * Parameters (fp32, BBRT_MLP_NPARAMS): W1[64][64] (inputs 51..63 zero), b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], W4[16][64] (rows 3..15 zero),
 * b4[16] (3..15 zero).  Feature order: ray (6) then patch (45).
 * ------------------------------------------------------------------------ */
#define BBRT_MLP_NIN 51
#define BBRT_MLP_NPARAMS 13520
typedef struct bbrt_mlp_s* bbrt_mlp;

/* Glorot-uniform weights from SplitMix64(seed), zero biases. */
int bbrt_mlp_create(uint64_t seed, bbrt_mlp* out);
void bbrt_mlp_destroy(bbrt_mlp mlp);
int bbrt_mlp_get_params(bbrt_mlp mlp, float* params);
int bbrt_mlp_set_params(bbrt_mlp mlp, const float* params);   /* pads must be zero */
/* Forward pass on the GPU: x [n][51] -> y [n][3] (host arrays). */
int bbrt_mlp_forward(bbrt_mlp mlp, const float* x, int64_t n, float* y);
/* Mean squared error over the batch and its gradient [BBRT_MLP_NPARAMS], by
 * the training kernels (host arrays, n <= 1<<20). */
int bbrt_mlp_loss_grad(bbrt_mlp mlp, const float* x, const float* target, int64_t n, float* loss, float* grad);

typedef struct {
    int64_t n_samples;   /* (ray, patch) pairs drawn; pairs stage 1 misses are dropped */
    uint64_t seed;
    int32_t batch;       /* samples per Adam step, a multiple of 64 */
    int32_t epochs;
    float lr, lr_final;  /* cosine decay from lr to lr_final */
    float val_frac, test_frac;
} bbrt_train_config;
typedef struct {
    int64_t n_pairs;     /* solved pairs kept (train + validation + test) */
    int64_t n_train, n_val, n_test;
    int64_t steps;
    float first_loss, final_loss;   /* training loss of the first and last step */
    float val_loss, test_loss;
    double gen_ms, train_ms;
} bbrt_train_report;
/* Draws cfg->n_samples pairs from `scene` (patch uniform, (u, v) uniform in
 * [-1, 1]^2, direction uniform over the front hemisphere, origin 0.25..4
 * patch scales back along it), solves each with stage 1 + converged Newton,
 * and trains.  loss_curve (may be NULL) receives one value per step. */
int bbrt_mlp_train(bbrt_mlp mlp, bbrt_scene scene, const bbrt_train_config* cfg, float* loss_curve,
                   bbrt_train_report* report);
/* Pairs as the trainer draws them (sample indices first .. first+n-1): patch,
 * origin, direction, features x [n][51] (fp32 values of the bf16 inputs' source),
 * target [n][3] = (u - u_a, v - v_a, (t - t_c)/scale) (see bbrt_mlp_features) and valid (stage 1 hit and
 * Newton converged). */
int bbrt_mlp_sample_pairs(bbrt_scene scene, uint64_t seed, int64_t first, int64_t n, int32_t* patch, double* origin,
                          double* dir, float* x, float* target, int8_t* valid);
/* Features of given (patch, ray) pairs: x [n][51] and base [n][3] = (t_c, u_a, v_a), the ray parameter of
 * the point closest to the patch centre and the clamped tangent-plane estimate of (u, v) that the network
 * corrects (seed = (u_a + y0, v_a + y1, t_c + scale y2)); frame [n_patches][24] (may be NULL): centre, e1,
 * e2, normal, scale, pad[3], the tangent-plane rows a_u and a_v, pad[2]. */
int bbrt_mlp_features(bbrt_scene scene, const int32_t* patch, const double* origin, const double* dir, int64_t n,
                      float* x, double* base, double* frame);
/* Attaches a snapshot of the network to the scene (per-patch frames and the
 * patch half of layer 1 are precomputed); BBRT_SEED_LEARNED then uses it.
 * Waits first for all work on the scene's device, so traces already enqueued
 * (bbrt_trace_device_async, bbrt_trace_device_noalloc, on any stream) finish
 * with the network they were enqueued with; traces enqueued after the call
 * returns use the new one. */
int bbrt_scene_set_mlp(bbrt_scene scene, bbrt_mlp mlp);
/* Seed A/B on sampled pairs (indices first .. first+n-1, same draw as
 * training): for each pair, stage 1, then converged Newton from the triangle
 * seed and from the learned seed.  Per pair: status [2][n] (1 hit, 0 miss,
 * -kind error), iters [2][n], accepted [n] (the learned result passes the
 * acceptance rule), seed_err [2][n] (max |du|, |dv| of each seed to the
 * triangle-seeded root), point_diff [n] (|p_learned - p_triangle|). */
int bbrt_mlp_seed_ab(bbrt_scene scene, uint64_t seed, int64_t first, int64_t n, int8_t* status, uint8_t* iters,
                     uint8_t* accepted, float* seed_err, double* point_diff);

/* Message of the last failed call on this thread ("" if none). */
const char* bbrt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BBRT_H */
