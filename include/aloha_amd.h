/*
 * aloha_amd.h -- C ABI of the MI355X-native per-brick render-and-composite
 * path.  Plain pointers and sizes only; no exceptions cross this boundary.
 *
 * Every entry point returns an int status: ALOHA_OK (0) or a negative
 * ALOHA_ERR_* code; aloha_last_error() gives the message.  Pointers named
 * d_* are device pointers (HBM); all others are host pointers.
 *
 * Each declaration names the reference interface it replaces
 * (vtvamr/aloha, file:line).
 */
#ifndef ALOHA_AMD_H
#define ALOHA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALOHA_OK 0
#define ALOHA_ERR_ARG (-1)
#define ALOHA_ERR_HIP (-2)
#define ALOHA_ERR_ALLOC (-3)
#define ALOHA_ERR_STATE (-4)
#define ALOHA_ERR_IO (-5)
#define ALOHA_ERR_LIMIT (-6)

/* voxel encodings (bytes per voxel, volume.h:76-86 / chopSuey -bpc) */
#define ALOHA_U8 1
#define ALOHA_U16 2
#define ALOHA_F32 4

/* jitter of the per-brick sample lattice */
#define ALOHA_JITTER_RANDOM 0 /* one RNG draw per brick visit (deviceCode.cpp:386) */
#define ALOHA_JITTER_FIXED 1  /* constant jitter_value (rank-invariant images) */

/* render flags */
#define ALOHA_FLAG_CROSSHAIR 1 /* invert row H/2 and column W/2 in rgba8 (deviceCode.cpp:410-414) */
#define ALOHA_FLAG_HEATMAP 2   /* blend a per-pixel clock heat map (deviceCode.cpp:399-403) */
#define ALOHA_FLAG_HEAT_CLOCK 4 /* with HEATMAP: every pixel's clock delta is params.heat_clock
                                   (instead of clock64() around its march), a reproducible
                                   heat map for checking the ramp and blend */

/* limits of one render launch (kernel-argument tables) */
#define ALOHA_MAX_RANKS 64
#define ALOHA_MAX_RANK_BRICKS 256
#define ALOHA_MAX_TF 4096

/* owl box3i / box3f / interval<float> layouts (2 x vec3 of 4-byte scalars) */
typedef struct { int32_t lower[3], upper[3]; } aloha_box3i;
typedef struct { float lower[3], upper[3]; } aloha_box3f;
typedef struct { float lower, upper; } aloha_range1f;

/* One brick record: the per-brick header of the .vol format
 * (chopSuey.cpp:342-345) and the VolBrick device record minus its texture
 * (deviceCode.h:35-50; brick.h:32-41).  80 bytes. */
typedef struct {
  aloha_box3i cell_range;
  aloha_box3i voxel_range;
  aloha_box3f space_range;
  aloha_range1f value_range;
} aloha_brick_desc;

/* Per-frame render settings: LaunchParams.render + accumID/island fields
 * (deviceCode.h:66-96), set by Renderer::render (waikikiRenderer.cpp:1021-1026).
 * dt is exposed here; the reference hard-codes 2.0 (waikikiRenderer.cpp:1023). */
typedef struct {
  float dt;
  float max_depth;
  int32_t spp;
  int32_t accum_id;
  int32_t island_id;
  int32_t num_islands;
  int32_t jitter_mode;
  float jitter_value;
  int32_t flags;
  float heat_map_scale;
  uint32_t heat_clock; /* clock delta of every pixel under ALOHA_FLAG_HEAT_CLOCK */
} aloha_render_params;

/* Per-rank outputs (device pointers, each may be NULL).  Rank slot s of a
 * launch writes pixel p at index s*W*H + p.
 *   rgba     float4, accum/(accumID+1)/spp before the crosshair (premultiplied)
 *   rgba8    make_rgba pack of the same (LaunchParams.fbPointer)
 *   depth    first brick entry / maxDepth, 1 if none (LaunchParams.fbDepth)
 *   accum    float4 running sum over frames (LaunchParams.accumBuffer)
 *   nsamples per-pixel executed sample count (measurement) */
typedef struct {
  float *rgba;
  uint32_t *rgba8;
  float *depth;
  float *accum;
  uint32_t *nsamples;
} aloha_images;

typedef struct aloha_ctx aloha_ctx;
typedef struct aloha_partitioner aloha_partitioner;

/* ---------------------------------------------------------------------- */
/* errors                                                                  */
/* ---------------------------------------------------------------------- */
/* message of the last failing call on ctx (ctx may be NULL: last failure of
 * a call that had no context). */
const char *aloha_last_error(const aloha_ctx *ctx);
/* library version string */
const char *aloha_version(void);
/* Test hook (no counterpart in the reference; for tests of the
 * allocation-failure paths): *made (may be NULL) = the device and pinned-host
 * allocations the library has made in this process so far; after >= 0 makes
 * the allocation `after` places further on fail as out of memory (the call
 * making it returns ALOHA_ERR_ALLOC and leaves its context as it was), once;
 * after < 0 disarms. */
int aloha_debug_alloc_fault(int64_t after, int64_t *made);

/* ---------------------------------------------------------------------- */
/* host-side bricking and partitioning (no GPU needed)                     */
/* ---------------------------------------------------------------------- */
/* VolumeSplitter (chopSuey.cpp:245-323): middle-split a dims volume into at
 * least num_bricks bricks (always splits once).  value_range is set to the
 * empty interval {FLT_MAX,-FLT_MAX}.  *count receives the brick count. */
int aloha_split_volume(const int32_t dims[3], int32_t num_bricks, aloha_brick_desc *out,
                       int32_t cap, int32_t *count);

/* Volume::get (volume.h:46-98): copy voxel sub-box [range.lower,range.upper)
 * of an x-fastest raw volume (bpc 1/2/4) to float and extend value_range. */
int aloha_volume_extract(const void *raw, const int32_t dims[3], int32_t bpc,
                         const aloha_box3i *range, float *out, aloha_range1f *value_range);

/* Partitioner (waikikiRenderer.cpp:126-338). */
int aloha_partitioner_create_kd(const aloha_box3f *domains, int32_t n, int32_t num_ranks,
                                aloha_partitioner **out); /* partitionKD :150-270 */
int aloha_partitioner_create_rr(int32_t n, int32_t num_ranks,
                                aloha_partitioner **out); /* partitionRoundRobin :137-148 */
void aloha_partitioner_destroy(aloha_partitioner *p);
/* perRank[rank] (:334) */
int aloha_partitioner_rank_bricks(const aloha_partitioner *p, int32_t rank, int32_t *ids,
                                  int32_t cap, int32_t *count);
/* computeCompositeOrder(reference) (:281-316) */
int aloha_partitioner_composite_order(const aloha_partitioner *p, const float reference[3],
                                      int32_t *order, int32_t cap, int32_t *count);
/* kd nodes as {axis, plane, child1, child2} (KDNode :318-322), 16 bytes each */
int aloha_partitioner_kd_nodes(const aloha_partitioner *p, void *nodes, int32_t cap,
                               int32_t *count);

/* ---------------------------------------------------------------------- */
/* .vol brick files (chopSuey.cpp:325-349; waikikiRenderer.cpp:596-665)    */
/* ---------------------------------------------------------------------- */
/* Write header {u64 n; box3i cellRange; box3i voxelRange; box3f spaceRange}
 * and per brick {desc; float voxels[]} (voxels[i] has the brick's voxel count). */
int aloha_vol_write(const char *path, const aloha_box3i *cell_range,
                    const aloha_box3i *voxel_range, const aloha_box3f *space_range,
                    int32_t n, const aloha_brick_desc *bricks, const float *const *voxels);
/* Read the header and brick records (no voxels); *n receives the brick count. */
int aloha_vol_read_header(const char *path, aloha_box3i *cell_range, aloha_box3i *voxel_range,
                          aloha_box3f *space_range, aloha_brick_desc *bricks, int32_t cap,
                          int32_t *n);
/* Read the voxels of brick i into out (float, voxel count of brick i). */
int aloha_vol_read_brick(const char *path, int32_t i, float *out, size_t out_count);

/* ---------------------------------------------------------------------- */
/* seeded synthetic inputs (bench / tests)                                 */
/* ---------------------------------------------------------------------- */
/* Sum of Gaussian blobs, centres U[0,dims), sigma U[smin,smax], amplitude
 * U[amin,amax]; normalised round(255*f/max f) to u8. */
int aloha_synth_blobs_u8(const int32_t dims[3], int32_t num_blobs, uint64_t seed, float smin,
                         float smax, float amin, float amax, uint8_t *out);
/* Same field normalised to [0,1] float. */
int aloha_synth_blobs_f32(const int32_t dims[3], int32_t num_blobs, uint64_t seed, float smin,
                          float smax, float amin, float amax, float *out);
/* Fractal Brownian motion of seeded gradient noise, rescaled to [0,1]. */
int aloha_synth_fbm_f32(const int32_t dims[3], int32_t octaves, float base_period,
                        uint64_t seed, float *out);
/* Piecewise-linear RGBA TF through num_points seeded control points
 * (first at 0, last at 1, sorted interior, RGBA U[0,1]); n entries. */
int aloha_synth_tf(int32_t n, int32_t num_points, uint64_t seed, float *out_rgba);
/* Orbit pinhole camera looking at the centre of bounds from dist_factor x
 * its diagonal, seeded azimuth/elevation; out = org, dir_00, dir_du, dir_dv
 * in the viewer's convention (viewer.cpp:148-158). */
int aloha_synth_camera(const aloha_box3f *bounds, int32_t W, int32_t H, float fovy_deg,
                       float dist_factor, uint64_t seed, float out[12]);

/* ---------------------------------------------------------------------- */
/* device context (replaces the OWL context/params/raygen, waikikiRenderer */
/* .cpp:377-384, and the CUDA texture objects, :697-777, :945-1013)        */
/* ---------------------------------------------------------------------- */
int aloha_ctx_create(int32_t device, aloha_ctx **out);
int aloha_ctx_destroy(aloha_ctx *ctx);
/* Launch all work of ctx on this hipStream_t from now on (NULL: the null
 * stream).  A new context owns a non-blocking stream of its own. */
int aloha_set_stream(aloha_ctx *ctx, void *hip_stream);
int aloha_sync(aloha_ctx *ctx);

/* Brick upload (cudaMalloc3DArray+cudaMemcpy3D+texture, :697-777).  voxels
 * is x-fastest with the desc's voxel_range extent, encoding dtype
 * (ALOHA_U8/U16/F32), on the device if src_on_device.  Values are decoded
 * as Volume::get does (u8/255.f, u16/65535.f) and stored as float32, or,
 * after aloha_set_u8_storage(ctx, 1), u8 bricks are kept as bytes (a
 * quarter of the HBM) and decoded at sample time through a table of the
 * same floats -- the images are the same bit for bit either way.
 * *brick_id receives its id. */
int aloha_brick_upload(aloha_ctx *ctx, const aloha_brick_desc *desc, const void *voxels,
                       int32_t dtype, int32_t src_on_device, int32_t *brick_id);
/* Storage of u8 bricks uploaded from now on (no reference counterpart: the
 * reference's float textures hold every brick as float32): keep_u8 != 0
 * keeps them as bytes (for volumes that would not fit as float32; the
 * march kernels that decode them run at fewer waves per SIMD), 0 (the
 * default) widens them to float32, which the differentiable form needs. */
int aloha_set_u8_storage(aloha_ctx *ctx, int32_t keep_u8);
/* *dtype = how brick_id is stored in HBM: ALOHA_U8 or ALOHA_F32. */
int aloha_brick_dtype(const aloha_ctx *ctx, int32_t brick_id, int32_t *dtype);
int aloha_brick_count(const aloha_ctx *ctx, int32_t *count);
/* device bytes held by bricks (padded layout) */
int aloha_brick_bytes(const aloha_ctx *ctx, uint64_t *bytes);

/* Transfer function (handleColorMapEvent, :945-1013): n RGBA float entries. */
int aloha_tf_set(aloha_ctx *ctx, const float *rgba, int32_t n);

/* Camera (handleCameraEvent, :853-871): org, dir_00, dir_du, dir_dv. */
int aloha_camera_set(aloha_ctx *ctx, const float org[3], const float dir_00[3],
                     const float dir_du[3], const float dir_dv[3]);

/* The raygen launch (owlLaunch2D(rayGen,W,H), :1028; renderFrame,
 * deviceCode.cpp:340-416) for P virtual ranks at once.  Rank slot s
 * (0<=s<P) marches bricks brick_ids[rank_offsets[s] .. rank_offsets[s+1]).
 * Asynchronous on the ctx stream. */
int aloha_render_ranks(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                       int32_t P, const int32_t *rank_offsets, const int32_t *brick_ids,
                       const aloha_images *out);

/* Fused raygen + sort-last composite: for every pixel, each rank slot of
 * the same table is marched from transparent and blended 'over' the running
 * colour in order[] -- the same result, bit for bit, as aloha_render_ranks
 * (spp>=1, accum_id 0, no heat map / crosshair) followed by
 * aloha_composite_over, without the ranks' images in HBM.  Outputs (device,
 * W*H; each may be NULL): float RGBA, RGBA8, per-pixel sample count over all
 * ranks. */
int aloha_render_composite(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                           int32_t H, int32_t P, const int32_t *rank_offsets,
                           const int32_t *brick_ids, const int32_t *order, float *d_out_rgba,
                           uint32_t *d_out_rgba8, uint32_t *d_nsamples);

/* One frame of Renderer::render (waikikiRenderer.cpp:1015-1083): the
 * raygen launch for every rank slot into `images` (as aloha_render_ranks;
 * float RGBA required) and IceT's ordered blend of their float RGBA images
 * in order[] into the outputs (as aloha_composite_over_sparse over the
 * rectangles of aloha_rank_rects) -- the same results bit for bit, in two
 * launches: the per-rank launch also writes the outputs transparent outside
 * the rectangles' bounding box, and the composite covers only that box.
 * images NULL: the caller wants the frame only; the rank images are then
 * stored compact, each only inside its rectangle (IceT's sparse images), in
 * a workspace the context holds (grown when a frame needs more; accum_id 0
 * and no heat map), with the same frame as result. */
int aloha_render_frame(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                       int32_t P, const int32_t *rank_offsets, const int32_t *brick_ids,
                       const aloha_images *images, const int32_t *order, float *d_out_rgba,
                       uint32_t *d_out_rgba8);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The frame of aloha_render_frame (same outputs, bit for
 * bit) without caller-held rank images: each rank's float RGBA image is kept
 * only inside its pixel rectangle (aloha_rank_rects), in a workspace of
 * ALOHA_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends.  Allocates nothing.  ALOHA_ERR_LIMIT when the
 * rectangles' pixels x 16 bytes exceed the workspace; accum_id must be 0 and
 * the heat map off (ALOHA_ERR_ARG).  Every call on one GPU uses the same
 * workspace, which keeps the last call's contents: calls that overlap on one
 * GPU (other streams or contexts) race; calls on one stream do not. */
#define ALOHA_NOALLOC_SCRATCH_BYTES (64u << 20)
int aloha_render_frame_noalloc(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                               int32_t H, int32_t P, const int32_t *rank_offsets,
                               const int32_t *brick_ids, const int32_t *order, float *d_out_rgba,
                               uint32_t *d_out_rgba8);

/* Schedule of the per-rank (aloha_render_ranks) and fused
 * (aloha_render_composite) march launches: lanes per ray, 1, 2, 4 or 8.  With
 * n > 1 the n lanes of a group evaluate n consecutive samples of one ray at
 * once and composite them in order; every schedule gives the same images bit
 * for bit.  Default 1. */
int aloha_set_march_schedule(aloha_ctx *ctx, int32_t ranks_lanes, int32_t fused_lanes);
int aloha_get_march_schedule(const aloha_ctx *ctx, int32_t *ranks_lanes, int32_t *fused_lanes);
/* Time both launches for every lanes-per-ray choice with these arguments
 * (reps launches each, on the context stream, synchronous) and keep the
 * fastest.  The fused launch is timed when order and an output are given.
 * table (optional): 4 rows {lanes, ranks ms, fused ms}; *rows = 4. */
int aloha_autotune_march(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                         int32_t P, const int32_t *rank_offsets, const int32_t *brick_ids,
                         const int32_t *order, const aloha_images *images, float *d_out_rgba,
                         uint32_t *d_out_rgba8, int32_t reps, float *table, int32_t *rows);
/* Count distinct voxels read by the march of the same launch (measurement:
 * compulsory bytes).  Synchronous. */
int aloha_render_footprint(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                           int32_t H, int32_t P, const int32_t *rank_offsets,
                           const int32_t *brick_ids, uint64_t *touched_voxels);
/* The same count per brick id: touched[id] for every uploaded brick (cap >=
 * aloha_brick_count), so the compulsory bytes follow each brick's storage. */
int aloha_render_footprint_bricks(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                  int32_t H, int32_t P, const int32_t *rank_offsets,
                                  const int32_t *brick_ids, uint64_t *touched, int32_t cap);

/* Sort-last ordered 'over' of P premultiplied float RGBA rank images, front
 * to back in order[] (IceT ORDERED BLEND, :911-916, :1043-1077), onto
 * transparent black.  d_rank_rgba: P*W*H float4.  Outputs may be NULL. */
int aloha_composite_over(aloha_ctx *ctx, const float *d_rank_rgba, int32_t P,
                         const int32_t *order, int32_t W, int32_t H, float *d_out_rgba,
                         uint32_t *d_out_rgba8);
/* Sparse form of aloha_composite_over.  IceT composites sparse images (it
 * skips runs of empty pixels); here rects[4*s .. 4*s+3] = x0, y0, x1, y1
 * bounds the non-transparent pixels of rank s's image, and only those are
 * read.  Bit-identical to the dense form whenever every rank pixel outside
 * its rectangle is transparent black -- which aloha_render_ranks guarantees
 * for the rectangles of aloha_rank_rects. */
int aloha_composite_over_sparse(aloha_ctx *ctx, const float *d_rank_rgba, int32_t P,
                                const int32_t *order, const int32_t *rects, int32_t W,
                                int32_t H, float *d_out_rgba, uint32_t *d_out_rgba8);
/* Pixel rectangles (P x 4: x0, y0, x1, y1) outside of which
 * aloha_render_ranks with these arguments writes transparent black float
 * RGBA: each rank's bricks projected with the current camera (the whole
 * image for a rank whose bricks reach the camera plane, and for every rank
 * when the heat map is on or accum_id > 0).  No counterpart in the
 * reference (its ranks' images go to IceT as they are). */
int aloha_rank_rects(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                     int32_t P, const int32_t *rank_offsets, const int32_t *brick_ids,
                     int32_t *rects);
/* Host only: pixel rectangle (x0, y0, x1, y1) outside of which no primary
 * ray of the camera (org, dir_00, dir_du, dir_dv; generateRay,
 * deviceCode.cpp:94-107, including its 1e-5 direction clamp) reaches any of
 * the n boxes; the whole image when no bound follows (a box reaching the
 * camera plane).  aloha_rank_rects applies it to each rank's bricks.  No
 * counterpart in the reference. */
int aloha_project_boxes(const float cam[12], const aloha_box3f *boxes, int32_t n, int32_t W,
                        int32_t H, int32_t rect[4]);
/* Same over P RGBA8 rank images (the reference's ubyte path, :908, :915). */
int aloha_composite_over_rgba8(aloha_ctx *ctx, const uint32_t *d_rank_rgba8, int32_t P,
                               const int32_t *order, int32_t W, int32_t H,
                               uint32_t *d_out_rgba8);
/* Depth composite (IceT Z_BUFFER/REDUCE, :903-909): per pixel the rank with
 * the smallest depth wins, ties to the lower rank.  Outputs may be NULL. */
int aloha_composite_zmin(aloha_ctx *ctx, const uint32_t *d_rank_rgba8, const float *d_rank_depth,
                         int32_t P, int32_t W, int32_t H, uint32_t *d_out_rgba8,
                         float *d_out_depth);

/* ---------------------------------------------------------------------- */
/* batched form (no reference counterpart: many views of the same virtual  */
/* ranks -- renderFrame + the IceT blend per view -- in one launch)         */
/* ---------------------------------------------------------------------- */
typedef struct aloha_batch aloha_batch;
/* V views (cams: V x 12 floats, host), one rank table, per-view visibility
 * orders (V x P, host).  Outputs (device, V*H*W; each may be NULL): float
 * RGBA, RGBA8, sample counts -- each view's image equals its
 * aloha_render_composite frame bit for bit.  Same params limits as that call. */
int aloha_batch_create(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                       int32_t V, const float *cams, int32_t P, const int32_t *rank_offsets,
                       const int32_t *brick_ids, const int32_t *orders, float *d_out_rgba,
                       uint32_t *d_out_rgba8, uint32_t *d_nsamples, aloha_batch **out);
int aloha_batch_destroy(aloha_batch *batch);
/* Schedule: lanes_per_ray 1, 2, 4 or 8 (the fused frame's lane groups,
 * aloha_set_march_schedule); persistent_blocks 0 = one workgroup per active
 * tile, else that many workgroups pulling (view, tile) items from
 * per-XCD counters.  Images do not depend on it.  Drops a captured graph. */
int aloha_batch_set_schedule(aloha_batch *batch, int32_t lanes_per_ray, int32_t persistent_blocks);
int aloha_batch_get_schedule(const aloha_batch *batch, int32_t *lanes_per_ray,
                             int32_t *persistent_blocks, int32_t *captured);
/* Tile of 64 pixels a workgroup composites: tile_w 8 (8x8), 16 (16x4, the
 * default and the single fused frame's) or 32 (32x2).  Images do not depend
 * on it.  Synchronous; drops a captured graph. */
int aloha_batch_set_tile(aloha_batch *batch, int32_t tile_w);
int aloha_batch_get_tile(const aloha_batch *batch, int32_t *tile_w);
/* Time every schedule (3 tiles x 4 lanes x 4 block counts) over reps
 * launches, keep the fastest; table (optional, cap rows) receives
 * {tile_w, lanes_per_ray, persistent_blocks, ms} per candidate.
 * Synchronous. */
int aloha_batch_autotune(aloha_batch *batch, int32_t reps, float *best_ms, float *table,
                         int32_t cap, int32_t *n);
/* Capture one launch (counter reset + batch kernel) into a hipGraph. */
int aloha_batch_capture(aloha_batch *batch);
/* Run the batch (the graph if captured), ordered on the context stream. */
int aloha_batch_launch(aloha_batch *batch);

/* ---------------------------------------------------------------------- */
/* inputs from device memory (training loops update them every step)      */
/* ---------------------------------------------------------------------- */
/* TF from n device-resident RGBA float entries; asynchronous on the stream. */
int aloha_tf_set_device(aloha_ctx *ctx, const float *d_rgba, int32_t n);
/* Replace the voxels of an uploaded brick (same extent) from d_voxels
 * (x-fastest, encoding dtype; a brick kept as u8 takes u8 only);
 * asynchronous on the stream. */
int aloha_brick_update(aloha_ctx *ctx, int32_t brick_id, const void *d_voxels, int32_t dtype);

/* ---------------------------------------------------------------------- */
/* differentiable form (no reference counterpart: the adjoint of           */
/* integrateBrick, deviceCode.cpp:256-298, for one rank's bricks)          */
/* ---------------------------------------------------------------------- */
#define ALOHA_MAX_VIEWS 16
/* Forward: V views (cams: V x {org, dir_00, dir_du, dir_dv} = 12 floats, host)
 * of one rank marching bricks brick_ids[0..nb).  Samples on the lattice
 * t_k = dt*(j+k0+k) with opacity clamped to 0.9999 (so the march can be
 * reversed).  d_rgba: V*H*W float4; d_nsamples: V*H*W.  Uses params->dt,
 * jitter_mode, jitter_value; spp 1, accumID 0. */
int aloha_diff_forward(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                       int32_t V, const float *cams, int32_t nb, const int32_t *brick_ids,
                       float *d_rgba, uint32_t *d_nsamples);
/* Adjoint of aloha_diff_forward: given its outputs and d_dl_drgba (V*H*W
 * float4), ADD dLoss/dTF into d_dtf (n x 4 floats) and dLoss/dvoxel into
 * d_dvoxels[i] (device, brick i's unpadded voxel count, x-fastest; a NULL
 * entry skips that brick).  Transmittance is recovered backward from the
 * forward's final alpha (the forward keeps nothing per sample); the adjoint
 * re-evaluates the samples into a context workspace (88 B per sample slot,
 * grown on demand) and sums every gradient in 64-bit fixed point in a fixed
 * order, so two calls give the same bits.  Waits for the stream once (the
 * sample total sizes the workspace). */
int aloha_diff_backward(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                        int32_t V, const float *cams, int32_t nb, const int32_t *brick_ids,
                        const float *d_rgba, const uint32_t *d_nsamples,
                        const float *d_dl_drgba, float *d_dtf, float *const *d_dvoxels);

/* ---------------------------------------------------------------------- */
/* surface path: triangle meshes (Renderer::Mode::TriMesh)                */
/* ---------------------------------------------------------------------- */
typedef struct aloha_mesh aloha_mesh;
/* Indexed triangle mesh on the host: verts nverts x 3 floats, tris ntris x
 * 3 vertex ids (Mesh/Geometry, mesh.h:30-45).  Copies its inputs. */
int aloha_mesh_create(const float *verts, int64_t nverts, const int32_t *tris, int64_t ntris,
                      aloha_mesh **out);
void aloha_mesh_destroy(aloha_mesh *mesh);
int aloha_mesh_info(const aloha_mesh *mesh, int64_t *nverts, int64_t *ntris, aloha_box3f *bounds,
                    int32_t *nclusters);
/* Copy vertices / triangles out (triangles in cluster order once split). */
int aloha_mesh_get(const aloha_mesh *mesh, float *verts, int32_t *tris);
/* ranges: count x {first, last} triangle ranges; bounds: split domains. */
int aloha_mesh_clusters(const aloha_mesh *mesh, int64_t *ranges, aloha_box3f *bounds, int32_t cap,
                        int32_t *count);
/* Seeded test scene: num_spheres UV spheres (rings x segments, 2*segments*
 * (rings-1) triangles each), radius uniform [rmin_frac, rmax_frac] of the
 * largest domain extent, radially displaced by smooth noise of amplitude
 * disp_frac x radius, centres uniform inside the domain. */
int aloha_mesh_synth_spheres(int32_t num_spheres, int32_t rings, int32_t segments,
                             const aloha_box3f *domain, float rmin_frac, float rmax_frac,
                             float disp_frac, uint64_t seed, aloha_mesh **out);
/* MeshSplitter, median strategy (chopSuey.cpp:48-198): repeatedly split the
 * cluster with most triangles at the median primitive centroid of its
 * domain's longest axis, triangles with min vertex < plane to the left.
 * Reorders the mesh's triangles cluster by cluster. */
int aloha_mesh_split(aloha_mesh *mesh, int32_t num_clusters, int32_t *count);
/* Wavefront OBJ (Mesh::loadOBJ mesh.cpp:39-85 as chopSuey uses it,
 * chopSuey.cpp:400-421): all vertices, the triangulated faces of the first
 * shape (MeshSplitter reads geoms[0] only); bounds over all vertices. */
int aloha_mesh_read_obj(const char *path, aloha_mesh **out);
/* .tri files (MeshSplitter::saveTris chopSuey.cpp:200-218; reader
 * waikikiRenderer.cpp:394-451).  An unsplit mesh is written as one cluster. */
int aloha_tri_write(const char *path, const aloha_mesh *mesh);
int aloha_tri_read(const char *path, aloha_mesh **out);

/* Upload P ranks' triangle geometry (waikikiRenderer.cpp:508-593): rank r
 * owns clusters cluster_ids[rank_offsets[r] .. rank_offsets[r+1]), each
 * drawn in colour color_ids[k] (randomColor input, ModelCH).  Triangles are
 * flattened to three private vertices each and given a binned-SAH BVH (the
 * OptiX BLAS/TLAS build); primitive ids count the rank's triangles in list
 * order.  geom_ids receives P ids.  An unsplit mesh has one cluster, 0. */
int aloha_tris_upload(aloha_ctx *ctx, const aloha_mesh *mesh, int32_t P,
                      const int32_t *rank_offsets, const int32_t *cluster_ids,
                      const uint32_t *color_ids, int32_t *geom_ids);
/* The same flatten + BVH build on the host only (no context, no device):
 * one rank's clusters in the device layout aloha_tris_export returns.
 * nodes / tris may be NULL to query the sizes; root is INT32_MIN for an
 * empty rank. */
int aloha_tris_build_host(const aloha_mesh *mesh, int32_t n, const int32_t *cluster_ids,
                          const uint32_t *color_ids, int64_t *ntris, int64_t *nnodes,
                          int32_t *depth, void *nodes, float *tris, int32_t *root);
/* Triangle count, BVH internal nodes, BVH depth and device bytes of one geometry. */
int aloha_tris_info(const aloha_ctx *ctx, int32_t geom, int64_t *ntris, int64_t *nnodes,
                    int32_t *depth, uint64_t *bytes);
/* Host copy of one geometry's device layout (for checking): nodes nnodes x
 * 16 x 4 bytes, tris ntris x 12 floats (three float4 vertices whose w hold
 * the primitive id and colour id as integer bits), root reference. */
int aloha_tris_export(const aloha_ctx *ctx, int32_t geom, void *nodes, float *tris, int32_t *root);
/* Per-rank surface images (renderFrame with traceRay(world)+ModelCH,
 * deviceCode.cpp:340-416): rank r traces geometry geom_ids[r].  Uses the
 * context camera, params->max_depth/spp/accum_id/island_id/num_islands/
 * flags/heat_map_scale; images->rgba / rgba8 / depth / accum (device, P x H
 * x W) may each be NULL.  d_stats (may be NULL): P x H x W x 2 uint32 of
 * BVH nodes visited and triangles tested per pixel. */
int aloha_render_tris(aloha_ctx *ctx, const aloha_render_params *params, int32_t W, int32_t H,
                      int32_t P, const int32_t *geom_ids, const aloha_images *images,
                      uint32_t *d_stats);
/* Render + depth composite in one launch: equals aloha_render_tris (rgba8,
 * depth, accum_id 0) followed by aloha_composite_zmin bit for bit. */
int aloha_render_tris_composite(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                int32_t H, int32_t P, const int32_t *geom_ids,
                                uint32_t *d_out_rgba8, float *d_out_depth, uint32_t *d_stats);

#ifdef __cplusplus
}
#endif
#endif /* ALOHA_AMD_H */
