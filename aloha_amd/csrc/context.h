// context.h -- the device context record behind the C ABI, shared by the
// volume entry points (context.hip) and the surface entry points
// (context_tris.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_types.h"
#include "internal.h"

// One rank's triangle geometry in HBM: BVH nodes (64 B each) and triangles
// in BVH leaf order (three float4 vertices, 48 B each).
struct TriGeom {
  float4 *d_nodes = nullptr;
  float4 *d_tris = nullptr;
  int32_t root = 0, depth = 0;
  int64_t ntris = 0, nnodes = 0;
  uint64_t bytes = 0;
  aloha_box3f box{};  // the triangles' bounds widened like the node boxes
};

struct aloha_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<aloha::DevBrick> bricks;
  std::vector<void *> brick_mem;
  // u8 voxels uploaded from now on are kept as bytes (else widened to f32)
  bool u8_storage = false;
  std::vector<aloha_brick_desc> descs;
  aloha::DevBrick *d_bricks = nullptr;
  size_t d_bricks_cap = 0;
  bool bricks_dirty = true;
  uint64_t brick_bytes = 0;
  float4 *d_tf = nullptr;
  int tf_n = 0, tf_cap = 0;
  float cam[12] = {};
  bool has_cam = false;
  std::vector<TriGeom> geoms;
  // lanes per ray of the per-rank and fused march launches (1, 2, 4, 8)
  int march_lanes = 1, fused_lanes = 1;
  // persistent per-rank launch: waves per lanes-per-ray option, without and
  // with the u8 decode (and the dynamic LDS size they were computed for)
  int q_waves[8] = {};
  size_t q_waves_lds[8] = {};
  // bumped whenever the device address of the brick table or the TF, or the
  // TF's entry count, changes: a batch launch compares it with the value its
  // launch arguments were made with
  uint64_t gen = 0;
  // batches made on this context (detached when the context goes first)
  std::vector<aloha_batch *> batches;
  // this device's workspace of aloha_render_frame_noalloc (NULL if its
  // address could not be had: scratch_err says why)
  void *scratch = nullptr;
  std::string scratch_err;
  // workspace of aloha_diff_backward, grown on demand: bin counts, bounds
  // and offsets; ray offsets; the sample slots (resampled values, run
  // marks, records, bin lists); bin tiles; the sample total read back
  uint32_t *diff_bins = nullptr;
  size_t diff_bins_cap = 0;
  void *diff_rays = nullptr;
  size_t diff_rays_cap = 0;   // bytes
  float4 *diff_slots = nullptr;
  size_t diff_slots_cap = 0;  // slots
  float *diff_tiles = nullptr;
  size_t diff_tiles_cap = 0;  // floats
  double *diff_tf_part = nullptr;
  size_t diff_tf_part_cap = 0;  // doubles
  uint64_t *h_total = nullptr;  // pinned host words (the adjoint's slot counters)
  float4 *frame_ws = nullptr;   // compact rank images of aloha_render_frame without images
  size_t frame_ws_cap = 0;      // float4
  std::string err;
};

namespace aloha {

constexpr int kMaxDevices = 64;
// Device address of the noalloc workspace on `device` (current).
int scratch_on_device(int device, void **out);

// Launch arguments of the persistent per-rank kernel for march arguments
// `a`; sets waves to its grid for `lanes` lanes per ray.
int prepare_ranks(aloha_ctx *ctx, const MarchArgs &a, int lanes, RanksArgs &q, int &waves);

// The frame with compact rank images (aloha_render_frame without images,
// aloha_render_frame_noalloc), named `name` in its error messages.
int frame_compact(aloha_ctx *ctx, const char *name, bool noalloc,
                  const aloha_render_params *params, int32_t W, int32_t H, int32_t P,
                  const int32_t *rank_offsets, const int32_t *brick_ids, const int32_t *order,
                  float *d_out_rgba, uint32_t *d_out_rgba8);
// Record a message on ctx (and as the thread's global error); returns code.
int fail(aloha_ctx *ctx, int code, const char *fmt, ...);
// Every device and pinned-host allocation of the library goes through these
// (hipMalloc / hipHostMalloc), so a test can make the n-th one fail
// (aloha_debug_alloc_fault).
hipError_t dev_malloc(void **p, size_t bytes);
hipError_t host_malloc(void **p, size_t bytes);
// Validate the render parameters every render entry point shares.
int check_params(aloha_ctx *ctx, const aloha_render_params *p);

}  // namespace aloha

#define HIP_CHECK(ctx, call)                                                              \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::aloha::fail(ctx, e_ == hipErrorOutOfMemory ? ALOHA_ERR_ALLOC : ALOHA_ERR_HIP, \
                           "%s: %s (%s)", __func__, #call, hipGetErrorString(e_));        \
  } while (0)
