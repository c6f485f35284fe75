/* pm_capi.h -- C ABI of the MI355X depth pipeline (libpm_hip.so).
 *
 * Replaces the kernel-call layer of HKUST-Aerial-Robotics/probabilistic_mapping
 * (motion_stereo/{calc_cost,sgm,filter_cost,depth_fusion,convert}.h).  Those were C++-linkage
 * functions that took untyped device pointers (unsigned char*), read shapes from compile-time
 * constants and knobs from globals, synchronised the device, and called exit() on error.
 * Here:
 *   - every entry point is extern "C", takes typed device pointers plus explicit shapes,
 *     and returns a pm_status (never exits);
 *   - work is enqueued on the caller's stream (NULL = default stream) and is asynchronous;
 *   - stage entry points allocate nothing: the caller owns every buffer;
 *   - pm_last_error() returns a message for the last failure on the calling thread.
 *
 * Data layouts (device memory):
 *   image   : uint8 [H][W]                  (row-major, pitch = W)
 *   quad    : uint64 [H][W]  texel (x,y) holds four IEEE FP16 numbers, low half first:
 *             a = I(x,y), I(x+1,y) - a, I(x,y+1) - a, I(x+1,y+1) - I(x+1,y) - I(x,y+1) + a,
 *             with I = 0 outside the image (exact: all are integers in [-510, 510]) --
 *             one 8-byte load and three FMAs per bilinear sample
 *   cost    : [H][W][D] pixel-major, D contiguous (the reference's INDEX(y,x,d),
 *             motion_stereo/cuda_utils.h:34), element type per pm_cost_type
 *   agg     : float [H][W][D], same layout
 *   depth, filter state : float [H][W]
 */
#ifndef PM_CAPI_H
#define PM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pm_status {
  PM_OK = 0,
  PM_ERR_INVALID_ARGUMENT = 1,
  PM_ERR_OUT_OF_MEMORY = 2,
  PM_ERR_LAUNCH = 3,
  PM_ERR_INVALID_STATE = 4,
  PM_ERR_DEVICE = 5
} pm_status;

typedef enum pm_cost_type {
  PM_COST_F32 = 0,
  PM_COST_F16 = 1,
  PM_COST_BF16 = 2
} pm_cost_type;

typedef void* pm_stream; /* hipStream_t; NULL = the default stream */

/* Largest supported plane count and source count per keyframe. */
#define PM_MAX_PLANES 256
#define PM_MAX_SOURCES 16

typedef struct pm_shape {
  int32_t width, height, planes;
} pm_shape;

/* Plane-sweep homography of one measurement frame, FP64 setup rounded to float:
 * h = K R_r^T R_l K^-1 (row-major), t = K R_r^T (T_l - T_r).
 * Reference: StereoMapper::update, motion_stereo/stereo_mapper.cpp:79-90. */
typedef struct pm_homography {
  float h[9];
  float t[3];
} pm_homography;

/* SGM knobs (motion_stereo/parameters.cpp:10-14) plus the path count (4 = the reference's
 * L->R, R->L, T->B, B->T set, sgm.cu:119-131; 8 adds the four diagonals).  Penalties and
 * tau_so must be finite and non-negative. */
typedef struct pm_sgm_params {
  float pi1, pi2, tau_so, sgm_q1, sgm_q2;
  int32_t paths;
} pm_sgm_params;

typedef struct pm_camera {
  float fx, fy, cx, cy;
} pm_camera;

/* Gaussian x Beta filter state planes, each float [H][W]
 * (alpha0..sigma1 of motion_stereo/stereo_mapper.h:48-49). */
typedef struct pm_fusion_state {
  float *alpha, *beta, *mu, *sig;
} pm_fusion_state;

/* Message describing the last error on this thread ("" if none). */
const char* pm_last_error(void);

/* Library version string. */
const char* pm_version(void);

/* FP64 homography setup from camera->world poses (R row-major 3x3, T 3-vector) and K
 * (row-major 3x3), rounded to float as the reference passes it to ad_calc_cost.
 * Reference: motion_stereo/stereo_mapper.cpp:79-90. */
pm_status pm_homography_setup(const double K[9], const double R_l[9], const double T_l[3],
                              const double R_r[9], const double T_r[3], pm_homography* out);

/* Builds the quad texture of an image (border 0).  Replaces the texture binding of
 * ad_calc_cost (motion_stereo/calc_cost.cu:221-244) and the float upload of
 * StereoMapper::update (stereo_mapper.cpp:74-77). */
pm_status pm_make_quad(const uint8_t* image, uint64_t* quad, int32_t width, int32_t height,
                       pm_stream stream);

/* Plane-sweep matching cost fused over all measurement frames of one keyframe.
 * Replaces num_sources calls of ad_calc_cost (motion_stereo/calc_cost.h:13-20,
 * calc_cost.cu:11-253) on a volume reset to -1 (stereo_mapper.cpp:69-70).
 * idep of plane d = idep_min + d * dep_sample (the reference has idep_min = 0).
 * src_quads and homs are HOST arrays of num_sources entries (the quads are device buffers).
 * cost: device [H][W][D] of cost_type; -1 where no source saw the (pixel, plane).
 * width * height < 2^29 (texel byte offsets are 32-bit). */
pm_status pm_calc_cost(pm_shape shape, const uint8_t* ref, int32_t num_sources,
                       const uint64_t* const* src_quads, const pm_homography* homs,
                       float idep_min, float dep_sample, void* cost, pm_cost_type cost_type,
                       pm_stream stream);

/* Multi-path semi-global aggregation into agg (fully overwritten; no pre-zeroing needed).
 * Replaces sgm2 (motion_stereo/sgm.h:6-7, sgm.cu:5-135) on a zeroed output
 * (stereo_mapper.cpp:118-120). */
pm_status pm_sgm(pm_shape shape, const uint8_t* ref, const void* cost, pm_cost_type cost_type,
                 pm_sgm_params params, float* agg, pm_stream stream);

/* pm_sgm followed by pm_filter_cost on the same stream (the sgm2 + filter_cost pair of
 * StereoMapper::output, stereo_mapper.cpp:118-126). */
pm_status pm_sgm_filter(pm_shape shape, const uint8_t* ref, const void* cost,
                        pm_cost_type cost_type, pm_sgm_params params, float var_scale,
                        float idep_min, float dep_sample, float* agg, float* depth,
                        pm_stream stream);

/* pm_sgm_filter with every path in flight at once: the paths write their own L volumes to a
 * caller-provided workspace (pm_sgm_workspace_size bytes: paths * W * H * D floats, 16-byte
 * aligned) instead of updating agg one after another, and one pass sums them in path order
 * into agg and filters.  agg and depth are identical, bit for bit, to pm_sgm_filter's (the
 * sum runs in the same order); agg may be NULL when only depth is wanted. */
pm_status pm_sgm_workspace_size(pm_shape shape, int32_t paths, uint64_t* bytes);
pm_status pm_sgm_filter_paths(pm_shape shape, const uint8_t* ref, const void* cost,
                              pm_cost_type cost_type, pm_sgm_params params, float var_scale,
                              float idep_min, float dep_sample, void* workspace,
                              uint64_t workspace_bytes, float* agg, float* depth,
                              pm_stream stream);

/* Compact form of pm_sgm_filter_paths (no reference counterpart; the storage-precision
 * variant).  The L recurrence of every path runs exactly as in pm_sgm, but each path stores only
 * its excess over the cost, E = L - C' (0 <= E <= P2 on every step, sgm.cu:85-98), as 16-bit
 * fixed point with step max(pi2, pi2 * sgm_q2) / 65534 in the workspace
 * (pm_sgm_excess_workspace_size bytes: paths * W * H * D uint16, 16-byte aligned).  One pass
 * then rebuilds agg = paths * C + step * sum of the levels (exact integer sum, one fma; 0 at
 * invalid pixels, those with a negative cost on any plane, sgm.cu:33-38, 101-104), writes it
 * to agg (may be NULL) and filters it into depth.
 * |agg - sum_r L_r| <= paths * step / 2 + FP32 rounding. */
pm_status pm_sgm_excess_workspace_size(pm_shape shape, int32_t paths, uint64_t* bytes);
pm_status pm_sgm_filter_excess(pm_shape shape, const uint8_t* ref, const void* cost,
                               pm_cost_type cost_type, pm_sgm_params params, float var_scale,
                               float idep_min, float dep_sample, void* workspace,
                               uint64_t workspace_bytes, float* agg, float* depth,
                               pm_stream stream);

/* Winner-take-all, rejection rules and parabola subpixel fit -> depth [H][W].
 * Replaces filter_cost (motion_stereo/filter_cost.h:7-8, filter_cost.cu:3-70); depth is a
 * dense [H][W] map instead of a pitched one.  depth = 1 / (idep_min + s * dep_sample). */
pm_status pm_filter_cost(pm_shape shape, const float* agg, float var_scale, float idep_min,
                         float dep_sample, float* depth, pm_stream stream);

/* Winner-take-all over a cost volume without rejection rules or subpixel fit: the CASE 0
 * (raw cost) and CASE 1 (aggregate) outputs of StereoMapper::outputFusionCPU
 * (motion_stereo/stereo_mapper.cpp:628-706), computed on the GPU instead of after a
 * download.  Per pixel: m = argmin (lowest index on ties, starting from 100000);
 * min < FLOAT_EPS -> 10000, m at the first or last plane -> 0, else 1/(idep_min + m*dep_sample). */
pm_status pm_wta(pm_shape shape, const void* cost, pm_cost_type cost_type, float idep_min,
                 float dep_sample, float* depth, pm_stream stream);

/* Optional 3x3 median over valid depths (0 < z < 1000).  No reference counterpart: the
 * median stage of the 752x480 workload.  in and out must not alias. */
pm_status pm_median3x3(const float* in, float* out, int32_t width, int32_t height,
                       pm_stream stream);

/* One Gaussian x Beta fusion step.  Replaces depth_fuse (motion_stereo/depth_fusion.h:13-20,
 * depth_fusion.cu:8-280) and the convert calls around it (stereo_mapper.cpp:149-152,185-193).
 *   depth     : [H][W] filtered depth in
 *   fused     : [H][W] fused depth out; may be depth itself (in place, like tmpDepthMap)
 *   zbuf      : [H][W] uint64 z-buffer
 *   zbuf_next : NULL, or a second [H][W] uint64 buffer.  NULL: the call empties zbuf first
 *               (no initialisation needed).  Non-NULL: zbuf must be empty (every entry all
 *               bits set) on entry, and the call empties zbuf_next for the next step, which
 *               passes it as its zbuf (two buffers alternating, both emptied once with
 *               pm_zbuf_reset); this saves a launch per step.
 *   table     : optional [H][W] int32 output of the filled propagation table (NULL to skip)
 *   pre       : previous state; mu and sig of moved pixels are rewritten in place
 *   cur       : current state
 *   r_p2c, t_p2c : previous->current camera transform (stereo_mapper.cpp:155-156), row-major */
pm_status pm_depth_fuse(int32_t width, int32_t height, pm_camera cam, const float r_p2c[9],
                        const float t_p2c[3], const float* depth, float* fused, uint64_t* zbuf,
                        uint64_t* zbuf_next, int32_t* table, pm_fusion_state pre,
                        pm_fusion_state cur, pm_stream stream);
/* Empties n z-buffer entries (sets every bit), for pm_depth_fuse's alternating buffers. */
pm_status pm_zbuf_reset(uint64_t* zbuf, size_t n, pm_stream stream);

/* The two halves of pm_depth_fuse as separate calls, and a step that also starts the next
 * one.  No reference counterpart: depth_fuse (depth_fusion.cu:8-280) runs transformMap,
 * hole_filling, depth_predict and depthUpdate in one call.
 *
 * pm_depth_fuse_warp: the forward warp of a step alone (transformMap, depth_fusion.cu:8-78):
 * gates and transforms `state` with the previous->current pose, rewrites state.mu / state.sig
 * of moved pixels in place and inserts them into zbuf, which must be empty on entry.
 * zbuf_clear (optional, another buffer) is emptied for a later step.
 *
 * pm_depth_fuse_step: the rest of a step (hole fill, predict, update, depth_fusion.cu:80-225)
 * on the z-buffer a warp filled, with the options
 *   median    : non-zero: the measurement is the 3x3 median of `depth` (pm_median3x3's
 *               filter), computed in the same launch and written to depth_out if non-NULL;
 *               depth, depth_out and fused must then be three different buffers.  Zero:
 *               the measurement is depth itself (which may be fused, in place) and depth_out
 *               must be NULL.
 *   r_next, t_next : NULL, or the next step's previous->current pose: then the call also
 *               runs the next step's warp on the state it has just written (cur), in the same
 *               launch and with the same results as pm_depth_fuse_warp(cam, r_next, t_next,
 *               cur, zbuf_next, zbuf_clear) after it.  zbuf_next must be empty on entry;
 *               zbuf_clear (optional) is emptied for the step after next.  zbuf, zbuf_next and
 *               zbuf_clear are three different buffers: neighbouring pixels still read this
 *               step's zbuf while the launch runs.
 * A chain warp(1) -> step(1, next 2) -> step(2, next 3) -> ... -> step(n) over three
 * z-buffers in rotation (step i reads Z[i % 3], fills Z[(i + 1) % 3], empties Z[(i + 2) % 3])
 * gives the same fused depths, tables and states as pm_depth_fuse at every step, except that
 * a step's cur.mu / cur.sig already hold the warped values the next pm_depth_fuse would
 * write into them (the last step's state is untouched).  Callers that know each pose before
 * the previous step's output (a recorded sequence, a batch of keyframes) run one launch per
 * keyframe this way. */
pm_status pm_depth_fuse_warp(int32_t width, int32_t height, pm_camera cam, const float r_p2c[9],
                             const float t_p2c[3], pm_fusion_state state, uint64_t* zbuf,
                             uint64_t* zbuf_clear, pm_stream stream);
pm_status pm_depth_fuse_step(int32_t width, int32_t height, pm_camera cam, const float* depth,
                             int32_t median, float* depth_out, float* fused,
                             const uint64_t* zbuf, int32_t* table, pm_fusion_state pre,
                             pm_fusion_state cur, const float* r_next, const float* t_next,
                             uint64_t* zbuf_next, uint64_t* zbuf_clear, pm_stream stream);

/* pm_depth_fuse without a z-buffer argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The z-buffer is a library-owned device array of
 * PM_FUSE_NOALLOC_MAX_PIXELS entries (16 MiB, held by each GPU the library runs on until the
 * process ends); width * height above that limit is refused with PM_ERR_INVALID_ARGUMENT.
 * The call allocates nothing: the array's address is looked up on the first call on each
 * device, and each call empties the array with a fill kernel before using it.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (other streams, other threads)
 * race; serialise them.  Other arguments and results as pm_depth_fuse with zbuf_next = NULL. */
#define PM_FUSE_NOALLOC_MAX_PIXELS (1 << 21)
pm_status pm_depth_fuse_noalloc(int32_t width, int32_t height, pm_camera cam,
                                const float r_p2c[9], const float t_p2c[3], const float* depth,
                                float* fused, int32_t* table, pm_fusion_state pre,
                                pm_fusion_state cur, pm_stream stream);

/* The fusion step of outputFusionCPU with CASE >= 3 (stereo_mapper.cpp:719-800): the
 * reference's CPU rules (depthFuseCPU, :246-587) on the GPU -- propagation gate 0.5, a 7x7
 * hole fill, tau^2 = 64, initial variance 256, outlier density 1/16, beta + 1 on a missing
 * measurement, output gated on the posterior inlier ratio > 0.51 -- with the measurement
 * clamps (> 100: dropped and masked, > 10: dropped) and output masks (< 0.1 -> 0).  The
 * raster-order z-buffer and hole fill of the reference are reproduced exactly.
 *   depth  : [H][W] filtered depth in, the result out
 *   table  : optional [H][W] int32 output of the filled table, (y << 16) | x, 0 = empty
 *   work   : device scratch of pm_depth_fuse_cpu_rules_workspace_size bytes
 *   r_p2c, t_p2c : previous->current camera transform (:778-779), row-major, double */
pm_status pm_depth_fuse_cpu_rules_workspace_size(int32_t width, int32_t height,
                                                 uint64_t* bytes);
pm_status pm_depth_fuse_cpu_rules(int32_t width, int32_t height, pm_camera cam,
                                  const double r_p2c[9], const double t_p2c[3], float* depth,
                                  int32_t* table, void* work, uint64_t work_bytes,
                                  pm_fusion_state pre, pm_fusion_state cur, pm_stream stream);

/* Fills n floats with value (used for the state resets of outputFusion,
 * stereo_mapper.cpp:143-144). */
pm_status pm_fill_f32(float* dst, size_t n, float value, pm_stream stream);

/* ---- batched form (no reference counterpart: the reference runs one keyframe at a time) --
 * A batch covers up to 64 keyframes of any size, source count and plane count.  pm_batch_run
 * enqueues cost -> SGM -> filter for all of them with one cost launch, one launch per SGM
 * direction and plane class, and one filter launch per lane class, each with persistent
 * blocks striding over a work list; results equal the single-keyframe entry points'.
 * Fusion is sequential in time and stays with pm_depth_fuse. */
#define PM_MAX_BATCH 64

typedef struct pm_keyframe {
  int32_t width, height, planes, num_sources;
  float idep_min, dep_sample;
  const uint8_t* ref;                          /* device uint8 [H][W] */
  const uint64_t* src_quads[PM_MAX_SOURCES];   /* device quad textures */
  pm_homography homs[PM_MAX_SOURCES];
  void* cost;                                  /* device [H][W][D] of the batch cost type */
  float* agg;                                  /* device [H][W][D] */
  float* depth;                                /* device [H][W] */
} pm_keyframe;

/* Launch shapes (picked by pm_batch_autotune): cost tile rows 2, 4 or 8 with an LDS staging
 * chunk of 8 or 16 planes (4 or 8 with 8 rows; 2 rows only when no keyframe has more than 8
 * sources), SGM prefetch depth 8 or 16 steps, and persistent grids of blocks_per_cu x
 * (compute units) blocks (0 = one block per work item).  cost_plane_split 1 splits each cost
 * tile's planes into groups of 32 (keyframes whose plane count is a multiple of 16), 0 keeps
 * a tile whole.  The cost work list is divided over the chip's 8 XCDs by work (tiles x planes
 * x sources), each XCD taking one contiguous range.  sgm_rows 1 runs the COMPACT storage's
 * paths of keyframes with 16-bit costs and 32, 64, 96 or 128 planes four scanlines per wave
 * (16 lanes per scanline), 0 one scanline per wave; both give the same bits. */
typedef struct pm_batch_tuning {
  int32_t cost_tile_rows;
  int32_t cost_stage_planes;
  int32_t sgm_prefetch;
  int32_t cost_blocks_per_cu;
  int32_t sgm_blocks_per_cu;
  int32_t cost_plane_split;
  int32_t sgm_rows;
} pm_batch_tuning;

typedef struct pm_batch pm_batch;

/* SGM storage of a batch.  EXACT: the paths update agg one after another, as pm_sgm (results
 * equal pm_sgm + pm_filter_cost, bit for bit).  COMPACT: every path of every keyframe in one
 * launch, storing 16-bit excess volumes in a batch-owned scratch (paths * W * H * D uint16
 * per keyframe), then one combine + filter launch per lane class (results equal
 * pm_sgm_filter_excess, bit for bit; a keyframe's agg may then be NULL).  SWEEP: the
 * horizontal paths as COMPACT; the downward paths advance together row by row and store one
 * 16-bit volume of their summed excess (level step = downward paths x the excess step); the
 * upward paths do the same and rebuild the aggregate in registers, then filter on the spot
 * (3 W H D uint16 + 6 W D floats of scratch per keyframe; at most 128 planes; agg, if given,
 * receives the rebuilt aggregate). */
typedef enum pm_sgm_storage {
  PM_SGM_EXACT = 0,
  PM_SGM_COMPACT = 1,
  PM_SGM_SWEEP = 2
} pm_sgm_storage;

/* Copies the keyframe table and work lists to the device (allocates; call once).  The batch
 * works on its keyframes largest first (W H D, then source count); each keyframe's results go
 * to its own buffers, so the caller's order only names keyframes in error messages. */
pm_status pm_batch_create(const pm_keyframe* kfs, int32_t count, pm_cost_type cost_type,
                          pm_sgm_params params, float var_scale, pm_sgm_storage storage,
                          pm_batch** out);
void pm_batch_destroy(pm_batch* b);
pm_status pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t);
pm_status pm_batch_get_tuning(const pm_batch* b, pm_batch_tuning* t);
/* Enqueues cost, SGM and filter for every keyframe on stream. */
pm_status pm_batch_run(pm_batch* b, pm_stream stream);
/* One stage of pm_batch_run: 0 = cost, 1 = SGM (all paths), 2 = filter (COMPACT: combine +
 * filter). */
pm_status pm_batch_run_stage(pm_batch* b, int32_t stage, pm_stream stream);
/* Times the candidate launch shapes on this batch and keeps the fastest (overwrites outputs). */
pm_status pm_batch_autotune(pm_batch* b, pm_stream stream, int32_t reps, pm_batch_tuning* best);
/* Records pm_batch_run into a hipGraph (stream must not be the default stream) ... */
pm_status pm_batch_capture(pm_batch* b, pm_stream stream);
/* ... and replays it. */
pm_status pm_batch_launch(pm_batch* b, pm_stream stream);

/* ---- Upload-once frame ring -----------------------------------------------------------
 * Replaces the per-use upload and u8 -> float conversion of StereoMapper::update
 * (motion_stereo/stereo_mapper.cpp:73-77), which the ROS node repeats for every use of a
 * frame (evaluate.cc:209-233).  A ring holds the newest `capacity` frames on the device, each
 * uploaded once through pinned staging and converted once into its quad texture (as
 * pm_make_quad).  Frame ids are >= 0; frame f occupies slot f % capacity until another frame
 * is put there.  put enqueues on `stream` and records an event that readers wait for. */
typedef struct pm_frame_ring pm_frame_ring;
pm_status pm_frame_ring_create(int32_t width, int32_t height, int32_t capacity,
                               pm_frame_ring** out);
void pm_frame_ring_destroy(pm_frame_ring* r);
/* image: host u8 [height][width]; reusable as soon as put returns. */
pm_status pm_frame_ring_put(pm_frame_ring* r, int64_t frame, const uint8_t* image,
                            pm_stream stream);
/* Device pointers of a held frame (either may be NULL); PM_ERR_INVALID_STATE if the frame
 * was never put or has been evicted. */
pm_status pm_frame_ring_frame(const pm_frame_ring* r, int64_t frame, const uint8_t** image,
                              const uint64_t** quad);
/* Frames uploaded so far. */
pm_status pm_frame_ring_puts(const pm_frame_ring* r, int64_t* puts);

/* ---- StereoMapper handle (motion_stereo/stereo_mapper.h:26-63) ------------------------
 * The C++ class lives in include/pm/stereo_mapper.h; these wrappers expose it to C and
 * Python.  Images and results are HOST arrays; the mapper owns its device buffers. */
typedef struct pm_mapper pm_mapper;

pm_status pm_mapper_create(int32_t width, int32_t height, int32_t planes, pm_mapper** out);
void pm_mapper_destroy(pm_mapper* m);
/* StereoMapper::initIntrinsic, stereo_mapper.cpp:51-61 (bf is unused, as in the reference) */
pm_status pm_mapper_init_intrinsic(pm_mapper* m, const double K[9], float bf, float dep_sample);
/* StereoMapper::initReference, stereo_mapper.cpp:63-71 */
pm_status pm_mapper_init_reference(pm_mapper* m, const uint8_t* image);
/* StereoMapper::update, stereo_mapper.cpp:73-91 */
pm_status pm_mapper_update(pm_mapper* m, const uint8_t* image, const double R_l[9],
                           const double T_l[3], const double R_r[9], const double T_r[3]);
/* initReference / update reading frame `frame` of a ring (same results as passing its host
 * image; no upload).  Both only enqueue work on the mapper's stream, so a frame passed to
 * either must stay in the ring until the next output call returns: the output call fails with
 * PM_ERR_INVALID_STATE if a staged source was evicted, and a put into the reference's slot
 * on a stream not ordered after the mapper's races with its copy. */
pm_status pm_mapper_init_reference_frame(pm_mapper* m, const pm_frame_ring* ring,
                                         int64_t frame);
pm_status pm_mapper_update_frame(pm_mapper* m, const pm_frame_ring* ring, int64_t frame,
                                 const double R_l[9], const double T_l[3],
                                 const double R_r[9], const double T_r[3]);
/* StereoMapper::output, stereo_mapper.cpp:93-109; result: host float [H][W] */
pm_status pm_mapper_output(pm_mapper* m, float* result);
/* StereoMapper::outputFusion, stereo_mapper.cpp:111-244; result: host float [H][W] holds the
 * filtered (un-fused) depth, as in the reference; the fused depth is the "fuse_depth" buffer. */
pm_status pm_mapper_output_fusion(pm_mapper* m, float* result, const double R[9],
                                  const double t[3], float fx, float fy, float cx, float cy);
/* StereoMapper::outputFusionCPU CASE modes 0-2, stereo_mapper.cpp:625-719 (WTA on the raw
 * cost, WTA on the aggregate, filtered depth); result: host float [H][W]. */
pm_status pm_mapper_output_case(pm_mapper* m, int32_t case_mode, float* result);
/* StereoMapper::outputFusionCPU, stereo_mapper.cpp:625-800, dispatching on the CASE knob:
 * modes 0-2 as pm_mapper_output_case; CASE >= 3 runs the reference's CPU fusion rules
 * (pm_depth_fuse_cpu_rules) on the mapper's own state, and result: host float [H][W] is the
 * fused depth (zeros on the first call, which only resets). */
pm_status pm_mapper_output_fusion_cpu(pm_mapper* m, float* result, const double R[9],
                                      const double t[3], float fx, float fy, float cx, float cy);
/* Copies a member buffer to host memory.  Names: "raw_cost", "sgm_cost" (float [H][W][D],
 * raw_cost converted from its storage type), "dep", "fuse_depth", "table" (int32),
 * "alpha", "beta", "mu", "sig" (the current state after the last fusion step),
 * "cpu_alpha", "cpu_beta", "cpu_mu", "cpu_sig", "cpu_table" (the same for outputFusionCPU's
 * CASE >= 3 step; the table is (y << 16) | x),
 * "img_l" (uint8).  bytes must equal the buffer size. */
pm_status pm_mapper_download(pm_mapper* m, const char* name, void* host, size_t bytes);
/* Storage type of the cost volume (default PM_COST_F32).  Takes effect at the next
 * initReference. */
pm_status pm_mapper_set_cost_type(pm_mapper* m, pm_cost_type type);

/* ---- keyframe / source selection driver (evaluate.cc:137-276, processImg) ---------------
 * The reference's ROS node without ROS (C++: include/pm/frame_driver.h).  Images and poses are
 * grabbed in arrival order (GrabImage / GrabPose, evaluate.cc:75-110); each process call is
 * one processImg step: the newest image as the reference, its pose by exact timestamp, and
 * agg_num measurement frames every interval-th image further back, with the poses as many
 * places further down the pose list; then StereoMapper initReference / update /
 * outputFusionCPU (CASE as set by pm_set_knobs). */
typedef struct pm_driver pm_driver;
typedef struct pm_driver_config {
  int32_t agg_num;    /* aggNum, evaluate.cc:24 (3) */
  int32_t interval;   /* interval, evaluate.cc:25 (6) */
  int32_t planes;     /* DEP_CNT (64) */
  double fx, fy, cx, cy;   /* evaluate.cc:37-40 */
  float dep_sample;   /* evaluate.cc:347: 1 / (0.15 * 160) */
  int32_t frame_ring; /* 1: frames reach the mapper through an upload-once frame ring of
                         agg_num * interval + 1 frames; 0: uploaded at every use, as the
                         reference does.  The depths are identical. */
} pm_driver_config;
/* with_mapper = 0 only plans the steps (host only: no device memory, no GPU work). */
pm_status pm_driver_create(const pm_driver_config* cfg, int32_t with_mapper, pm_driver** out);
void pm_driver_destroy(pm_driver* d);
/* GrabImage: an 8-bit grayscale width x height host image (copied). */
pm_status pm_driver_grab_image(pm_driver* d, int64_t stamp_ns, const uint8_t* image,
                               int32_t width, int32_t height);
/* GrabPose: pose = (tx, ty, tz, qx, qy, qz, qw), camera -> world. */
pm_status pm_driver_grab_pose(pm_driver* d, int64_t stamp_ns, const double pose[7]);
/* One processImg step.  *ran = 1 if it ran; then stamp_ns, ref_image and src_images[agg_num]
 * (grab-order indices, each may be NULL) describe it, and depth (host float [H][W], NULL for
 * a planning driver) receives the 32FC1 result. */
pm_status pm_driver_process(pm_driver* d, int32_t* ran, int64_t* stamp_ns, int32_t* ref_image,
                            int32_t* src_images, float* depth, size_t depth_bytes);
/* Steps not run because the sources' poses were not buffered yet (the reference reads past
 * the end of its pose list there). */
pm_status pm_driver_skipped(const pm_driver* d, int32_t* skipped);
/* Buffered message counts (the lists grow by about one message per step, as the
 * reference's do) and how many buffered images still hold their pixels (at most
 * agg_num * interval + 1: older ones can never be read again).  Each pointer may be NULL. */
pm_status pm_driver_buffered(const pm_driver* d, int32_t* images, int32_t* poses,
                             int32_t* images_with_pixels);
/* Frames uploaded through the driver's frame ring so far (0 without a ring). */
pm_status pm_driver_ring_puts(const pm_driver* d, int64_t* puts);
/* Eigen's Quaterniond(w, x, y, z).toRotationMatrix(), q = (x, y, z, w); R row-major. */
pm_status pm_quaternion_to_rotation(const double q[4], double R[9]);

/* ---- knobs (motion_stereo/parameters.cpp:10-18, plus additive ones) ------------------- */
typedef struct pm_knobs {
  float pi1, pi2, tau_so, sgm_q1, sgm_q2, var_scale;
  int32_t case_mode;   /* CASE */
  float idep_min;      /* additive: first plane's inverse depth (reference: 0) */
  int32_t sgm_paths;   /* additive: 4 (reference) or 8 */
  int32_t median;      /* additive: 1 = 3x3 median after the filter (reference: 0) */
} pm_knobs;
void pm_get_knobs(pm_knobs* out);
void pm_set_knobs(const pm_knobs* in);

/* ---- differentiable form: soft-argmin depth readout and its adjoint ------------------------
 * No counterpart in the reference.  Forward: the reference's cost (calc_cost.cu:11-210) on float
 * reference intensities, the multi-path SGM (sgm.cu:5-113), then
 *   depth_hat(p) = sum_d softmax_d(-agg(p,d) / (cost_scale * temperature)) / idep_d
 *   loss = mean over M of |depth_hat - gt|,  M = pixels with a valid cost and gt > 0.
 * Backward: dL/dI_ref (float [H][W]) and dL/d(H, t) per source (double [S][12]: h11..h33,
 * t1..t3 of pm_homography), from a hand-written adjoint; pm_se3_pose_grad maps the latter to
 * the se3 pose.  The handle owns its device buffers (cost, source counts, aggregate, and
 * either all per-path L volumes or, with recompute = 1, one). */
typedef struct pm_diff_params {
  float temperature;   /* softmax temperature T */
  float cost_scale;    /* aggregate cost units per logit unit at T = 1 (255: full-scale intensity) */
  int32_t recompute;   /* 1: the backward recomputes the aggregate and the per-path L volumes */
} pm_diff_params;

typedef struct pm_diff pm_diff;

pm_status pm_diff_create(pm_shape shape, int32_t num_sources, pm_sgm_params sgm,
                         pm_diff_params params, pm_diff** out);
void pm_diff_destroy(pm_diff* h);
/* Device bytes held by the handle. */
pm_status pm_diff_memory(const pm_diff* h, uint64_t* bytes);
/* ref: device float [H][W]; src_quads: host array of num_sources device quad textures;
 * homs: host array (FP32 H, t per source); gt: device float [H][W] (<= 0 or non-finite =
 * unknown); depth: device float [H][W] (depth_hat); loss: device double.  ref, the quads, gt
 * and depth must stay valid until pm_diff_backward has run. */
pm_status pm_diff_forward(pm_diff* h, const float* ref, const uint64_t* const* src_quads,
                          const pm_homography* homs, float idep_min, float dep_sample,
                          const float* gt, float* depth, double* loss, pm_stream stream);
/* grad_ref: device float [H][W]; grad_homs: device double [S][12]; both overwritten.  With
 * recompute = 0 the backward consumes the forward's stored volumes (the aggregate becomes dL/dagg
 * and the per-path L volumes the per-path adjoints): a second backward needs another forward
 * (PM_ERR_INVALID_STATE otherwise).  With recompute = 1 it recomputes them and may run again. */
pm_status pm_diff_backward(pm_diff* h, float grad_loss, float* grad_ref, double* grad_homs,
                           pm_stream stream);

/* Copies a handle buffer to host memory (synchronous): "cost" (float [H][W][D]), "count"
 * (uint8 sources per (pixel, plane)), "agg" (float: the aggregate after the forward, dL/dagg
 * after the backward), "grad_cost" (float dL/dC after the backward), "valid" (uint8 [H][W]).
 * bytes must equal the buffer size. */
pm_status pm_diff_download(const pm_diff* h, const char* name, void* host, size_t bytes);

/* se3 pose of a measurement frame (host, FP64).  xi = (rho, phi), translation first;
 * (R, t) = Exp(xi) (R0, t0) takes reference-camera coordinates to the measurement camera.
 * pm_se3_homography: H = K R K^-1, t_h = K t (the relative form of stereo_mapper.cpp:79-80).
 * pm_se3_pose_grad: dL/dxi from dL/d(H, t_h) (grad_hom[12] as in pm_diff_backward). */
pm_status pm_se3_homography(const double K[9], const double R0[9], const double t0[3],
                            const double xi[6], double H[9], double t[3]);
pm_status pm_se3_pose_grad(const double K[9], const double R0[9], const double t0[3],
                           const double xi[6], const double grad_hom[12], double grad_xi[6]);

#ifdef __cplusplus
}
#endif

#endif /* PM_CAPI_H */
