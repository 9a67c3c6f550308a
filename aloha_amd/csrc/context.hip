// context.hip -- the device context behind the C ABI: brick storage in HBM,
// transfer function, camera, and the render / composite entry points.
// Replaces the OWL context, launch params and CUDA texture objects that
// Renderer builds (waikikiRenderer.cpp:377-384, :697-797, :853-1013) and the
// IceT composite it runs per frame (:1030-1079).
//
// Every entry point validates its arguments on the host and returns an
// ALOHA_ERR_* code with a message instead of launching anything it cannot
// prove in bounds.  Render and composite calls only enqueue kernels on the
// context's stream (no allocation, no synchronisation), so a caller may
// capture them in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "context.h"

using namespace aloha;

namespace aloha {

namespace {
std::atomic<int64_t> g_allocs{0};    // allocations made through dev_malloc / host_malloc
std::atomic<int64_t> g_fail_at{-1};  // index of the allocation that fails (-1: none)

bool alloc_fault() {
  const int64_t k = g_allocs.fetch_add(1);
  int64_t at = k;
  // one-shot: the armed allocation fails and disarms the fault
  return g_fail_at.compare_exchange_strong(at, -1);
}
}  // namespace

int64_t g_allocs_total() { return g_allocs.load(); }

hipError_t dev_malloc(void **p, size_t bytes) {
  *p = nullptr;
  if (alloc_fault()) return hipErrorOutOfMemory;
  return hipMalloc(p, bytes);
}

hipError_t host_malloc(void **p, size_t bytes) {
  *p = nullptr;
  if (alloc_fault()) return hipErrorOutOfMemory;
  return hipHostMalloc(p, bytes);
}

int fail(aloha_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  global_error() = buf;
  return code;
}

}  // namespace aloha

namespace {

bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

int sync_bricks(aloha_ctx *ctx) {
  if (!ctx->bricks_dirty) return ALOHA_OK;
  if (ctx->bricks.size() > ctx->d_bricks_cap) {
    if (ctx->d_bricks) HIP_CHECK(ctx, hipFree(ctx->d_bricks));
    ctx->d_bricks = nullptr;
    ctx->d_bricks_cap = 0;  // a failed allocation leaves no table (and grows again next time)
    const size_t cap = std::max<size_t>(16, 2 * ctx->bricks.size());
    HIP_CHECK(ctx, dev_malloc((void **)&ctx->d_bricks, cap * sizeof(DevBrick)));
    ctx->d_bricks_cap = cap;
    ++ctx->gen;
  }
  if (!ctx->bricks.empty())
    HIP_CHECK(ctx, hipMemcpy(ctx->d_bricks, ctx->bricks.data(),
                             ctx->bricks.size() * sizeof(DevBrick), hipMemcpyHostToDevice));
  ctx->bricks_dirty = false;
  return ALOHA_OK;
}

}  // namespace

int aloha::check_params(aloha_ctx *ctx, const aloha_render_params *p) {
  if (!p) return fail(ctx, ALOHA_ERR_ARG, "render: params is NULL");
  if (!(p->dt > 0.f) || !std::isfinite(p->dt))
    return fail(ctx, ALOHA_ERR_ARG, "render: dt must be finite and > 0 (got %g)", p->dt);
  if (!std::isfinite(p->max_depth) || p->max_depth == 0.f)
    return fail(ctx, ALOHA_ERR_ARG, "render: max_depth must be finite and non-zero");
  if (p->spp < 1) return fail(ctx, ALOHA_ERR_ARG, "render: spp must be >= 1");
  if (p->accum_id < 0) return fail(ctx, ALOHA_ERR_ARG, "render: accum_id must be >= 0");
  if (p->num_islands < 1 || p->island_id < 0 || p->island_id >= p->num_islands)
    return fail(ctx, ALOHA_ERR_ARG, "render: island %d of %d invalid", p->island_id,
                p->num_islands);
  if (p->jitter_mode != ALOHA_JITTER_RANDOM && p->jitter_mode != ALOHA_JITTER_FIXED)
    return fail(ctx, ALOHA_ERR_ARG, "render: unknown jitter_mode %d", p->jitter_mode);
  if (p->jitter_mode == ALOHA_JITTER_FIXED && !(p->jitter_value >= 0.f && p->jitter_value < 1.f))
    return fail(ctx, ALOHA_ERR_ARG, "render: jitter_value must be in [0,1)");
  return ALOHA_OK;
}

namespace {

// Pixel rectangle of the bricks bricks[ids[0..n)] (project_boxes over their
// space boxes): no primary ray of another pixel reaches them.
void rank_rect(const float cam[12], const std::vector<DevBrick> &bricks, const int32_t *ids,
               int n, int W, int H, int32_t out[4]) {
  std::vector<aloha_box3f> boxes((size_t)std::max(n, 0));
  for (int k = 0; k < n; ++k) {
    const DevBrick &b = bricks[ids[k]];
    for (int a = 0; a < 3; ++a) {
      boxes[k].lower[a] = b.space_lo[a];
      boxes[k].upper[a] = b.space_hi[a];
    }
  }
  project_boxes(cam, boxes.data(), n, W, H, out);
}

// The march advances t by repeated addition t += dt, so every t it reaches
// must satisfy t + dt > t, and one brick visit must stay a bounded loop:
// dt >= 2^-20 x the farthest brick corner's distance from org (t + dt then
// moves by many ulps), and a brick diagonal of at most 2^22 steps.
int check_steps(aloha_ctx *ctx, const float cam[12], const std::vector<DevBrick> &bricks,
                const int32_t *ids, int n, float dt) {
  for (int k = 0; k < n; ++k) {
    const DevBrick &b = bricks[ids[k]];
    double far2 = 0.0, diag2 = 0.0;
    for (int c = 0; c < 8; ++c) {
      double d2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double v = ((c >> a) & 1) ? b.space_hi[a] : b.space_lo[a];
        d2 += (v - cam[a]) * (v - cam[a]);
      }
      far2 = std::max(far2, d2);
    }
    for (int a = 0; a < 3; ++a)
      diag2 += ((double)b.space_hi[a] - b.space_lo[a]) * ((double)b.space_hi[a] - b.space_lo[a]);
    if (std::sqrt(far2) * 0x1p-20 > (double)dt || std::sqrt(diag2) > 0x1p22 * (double)dt)
      return fail(ctx, ALOHA_ERR_LIMIT,
                  "render: dt %g too small for brick %d (farthest corner at %g, diagonal %g)",
                  dt, ids[k], std::sqrt(far2), std::sqrt(diag2));
  }
  return ALOHA_OK;
}

// Fill MarchArgs from the context and a validated rank table.
int build_march(aloha_ctx *ctx, const aloha_render_params *p, int32_t W, int32_t H, int32_t P,
                const int32_t *rank_offsets, const int32_t *brick_ids, MarchArgs &a) {
  if (int rc = check_params(ctx, p)) return rc;
  if (W < 1 || H < 1 || (int64_t)W * H > (int64_t)1 << 30)
    return fail(ctx, ALOHA_ERR_ARG, "render: image %dx%d out of range", W, H);
  if (P < 1 || P > ALOHA_MAX_RANKS)
    return fail(ctx, ALOHA_ERR_LIMIT, "render: P=%d outside [1,%d]", P, ALOHA_MAX_RANKS);
  if (!rank_offsets || rank_offsets[0] != 0)
    return fail(ctx, ALOHA_ERR_ARG, "render: rank_offsets must start at 0");
  for (int s = 0; s < P; ++s)
    if (rank_offsets[s + 1] < rank_offsets[s])
      return fail(ctx, ALOHA_ERR_ARG, "render: rank_offsets not monotone at %d", s);
  const int nb = rank_offsets[P];
  if (nb > ALOHA_MAX_RANK_BRICKS)
    return fail(ctx, ALOHA_ERR_LIMIT, "render: %d brick slots exceed %d", nb,
                ALOHA_MAX_RANK_BRICKS);
  if (nb > 0 && !brick_ids) return fail(ctx, ALOHA_ERR_ARG, "render: brick_ids is NULL");
  for (int k = 0; k < nb; ++k)
    if (brick_ids[k] < 0 || brick_ids[k] >= (int)ctx->bricks.size())
      return fail(ctx, ALOHA_ERR_ARG, "render: brick id %d not uploaded", brick_ids[k]);
  if (!ctx->d_tf || ctx->tf_n < 1) return fail(ctx, ALOHA_ERR_STATE, "render: no transfer function set");
  if (!ctx->has_cam) return fail(ctx, ALOHA_ERR_STATE, "render: no camera set");
  if (int rc = check_steps(ctx, ctx->cam, ctx->bricks, brick_ids, nb, p->dt)) return rc;
  if (int rc = sync_bricks(ctx)) return rc;
  std::memset(&a, 0, sizeof a);
  std::memcpy(a.org, ctx->cam, sizeof a.org);
  std::memcpy(a.dir00, ctx->cam + 3, sizeof a.dir00);
  std::memcpy(a.du, ctx->cam + 6, sizeof a.du);
  std::memcpy(a.dv, ctx->cam + 9, sizeof a.dv);
  a.dt = p->dt;
  a.max_depth = p->max_depth;
  a.jitter_value = p->jitter_value;
  a.heat_scale = p->heat_map_scale;
  a.heat_clock = p->heat_clock;
  a.W = W;
  a.H = H;
  a.spp = p->spp;
  a.accum_id = p->accum_id;
  a.inv_frames = 1.f / (float)(p->accum_id + 1);
  a.inv_spp = 1.f / (float)p->spp;
  a.island_id = p->island_id;
  a.num_islands = p->num_islands;
  a.jitter_mode = p->jitter_mode;
  a.flags = p->flags;
  a.P = P;
  a.tf_n = ctx->tf_n;
  a.bricks = ctx->d_bricks;
  a.tf = ctx->d_tf;
  for (int s = 0; s <= P; ++s) a.rank_off[s] = rank_offsets[s];
  for (int k = 0; k < nb; ++k) {
    a.rank_bricks[k] = brick_ids[k];
    if (ctx->bricks[brick_ids[k]].dtype == ALOHA_U8) a.u8_lut = 1;
  }
  for (int s = 0; s < P; ++s)
    rank_rect(ctx->cam, ctx->bricks, brick_ids + rank_offsets[s],
              rank_offsets[s + 1] - rank_offsets[s], W, H, a.rect[s]);
  return ALOHA_OK;
}

}  // namespace

extern "C" const char *aloha_last_error(const aloha_ctx *ctx) {
  return ctx ? ctx->err.c_str() : global_error().c_str();
}

extern "C" int aloha_ctx_create(int32_t device, aloha_ctx **out) {
  if (!out) return fail(nullptr, ALOHA_ERR_ARG, "aloha_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1)
    return fail(nullptr, ALOHA_ERR_HIP, "aloha_ctx_create: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n)
    return fail(nullptr, ALOHA_ERR_ARG, "aloha_ctx_create: device %d of %d", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail(nullptr, ALOHA_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  auto *ctx = new aloha_ctx;
  ctx->device = device;
  // the workspace of aloha_render_frame_noalloc, looked up once per device
  if (hipError_t se = (hipError_t)scratch_on_device(device, &ctx->scratch); se != hipSuccess)
    ctx->scratch_err = hipGetErrorString(se);
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(nullptr, ALOHA_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  ctx->own_stream = true;
  *out = ctx;
  return ALOHA_OK;
}

namespace aloha {
void batch_detach(aloha_batch *bt);
}

extern "C" int aloha_ctx_destroy(aloha_ctx *ctx) {
  if (!ctx) return ALOHA_OK;
  hipSetDevice(ctx->device);
  // batches still alive lose their device resources now; destroying them
  // later only frees the host record
  for (aloha_batch *bt : ctx->batches) batch_detach(bt);
  ctx->batches.clear();
  hipStreamSynchronize(ctx->stream);
  for (void *m : ctx->brick_mem) hipFree(m);
  if (ctx->d_bricks) hipFree(ctx->d_bricks);
  if (ctx->d_tf) hipFree(ctx->d_tf);
  if (ctx->diff_bins) hipFree(ctx->diff_bins);
  if (ctx->diff_rays) hipFree(ctx->diff_rays);
  if (ctx->diff_slots) hipFree(ctx->diff_slots);
  if (ctx->diff_tiles) hipFree(ctx->diff_tiles);
  if (ctx->diff_tf_part) hipFree(ctx->diff_tf_part);
  if (ctx->frame_ws) hipFree(ctx->frame_ws);
  if (ctx->h_total) hipHostFree(ctx->h_total);
  for (TriGeom &g : ctx->geoms) {
    if (g.d_nodes) hipFree(g.d_nodes);
    if (g.d_tris) hipFree(g.d_tris);
  }
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return ALOHA_OK;
}

extern "C" int aloha_set_stream(aloha_ctx *ctx, void *hip_stream) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_set_stream: ctx is NULL");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (ctx->own_stream) {
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_CHECK(ctx, hipStreamDestroy(ctx->stream));
    ctx->own_stream = false;
    ctx->stream = nullptr;
  }
  // NULL is the device's default (null) stream, as hipStream_t 0 is in HIP
  ctx->stream = (hipStream_t)hip_stream;
  return ALOHA_OK;
}

extern "C" int aloha_sync(aloha_ctx *ctx) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_sync: ctx is NULL");
  HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_debug_alloc_fault(int64_t after, int64_t *made) {
  if (made) *made = g_allocs_total();
  g_fail_at.store(after < 0 ? -1 : g_allocs_total() + after);
  return ALOHA_OK;
}

extern "C" int aloha_brick_upload(aloha_ctx *ctx, const aloha_brick_desc *desc,
                                  const void *voxels, int32_t dtype, int32_t src_on_device,
                                  int32_t *brick_id) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_brick_upload: ctx is NULL");
  if (!desc || !voxels) return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_upload: NULL desc/voxels");
  if (dtype != ALOHA_U8 && dtype != ALOHA_U16 && dtype != ALOHA_F32)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_upload: unknown dtype %d", dtype);
  int32_t n[3];
  for (int a = 0; a < 3; ++a) {
    n[a] = desc->voxel_range.upper[a] - desc->voxel_range.lower[a];
    if (n[a] < 1)
      return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_upload: empty voxel range on axis %d", a);
    if (!(desc->space_range.lower[a] <= desc->space_range.upper[a]))
      return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_upload: inverted space range on axis %d", a);
  }
  if (!finite3(desc->space_range.lower) || !finite3(desc->space_range.upper))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_upload: non-finite space range");
  const int64_t px = (int64_t)n[0] + kPadLo + kPadHi;
  const int64_t py = (int64_t)n[1] + kPadLo + kPadHi;
  const int64_t pz = (int64_t)n[2] + kPadLo + kPadHi;
  const int64_t padded = px * py * pz;
  // offsets are 32-bit and the kernels form them with 24-bit multiplies
  if (padded >= ((int64_t)1 << 31) || px * py >= ((int64_t)1 << 24))
    return fail(ctx, ALOHA_ERR_LIMIT,
                "aloha_brick_upload: brick of %lld padded voxels (slice %lld) exceeds 2^31 (2^24)",
                (long long)padded, (long long)(px * py));
  const size_t src_bytes = (size_t)n[0] * n[1] * n[2] * (size_t)dtype;
  // u8 voxels stay bytes in HBM when the context keeps them (a quarter of
  // the f32 footprint; the kernels decode through u/255.f), else f32
  const int store = dtype == ALOHA_U8 && ctx->u8_storage ? ALOHA_U8 : ALOHA_F32;
  const uint64_t mem_bytes =
      store == ALOHA_U8 ? (uint64_t)padded + kU8Slack : (uint64_t)padded * sizeof(float);
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  void *mem = nullptr;
  HIP_CHECK(ctx, dev_malloc(&mem, mem_bytes));
  const void *src = voxels;
  void *staging = nullptr;
  if (!src_on_device) {
    hipError_t e = dev_malloc((void **)&staging, src_bytes);
    if (e == hipSuccess) e = hipMemcpy(staging, voxels, src_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(mem);
      if (staging) hipFree(staging);
      return fail(ctx, e == hipErrorOutOfMemory ? ALOHA_ERR_ALLOC : ALOHA_ERR_HIP,
                  "aloha_brick_upload: staging copy: %s", hipGetErrorString(e));
    }
    src = staging;
  }
  hipError_t e = launch_pad_brick(src, dtype, n, mem, store, (int32_t)px, (int32_t)py, (int32_t)pz,
                                  ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (staging) hipFree(staging);
  if (e != hipSuccess) {
    hipFree(mem);
    return fail(ctx, ALOHA_ERR_HIP, "aloha_brick_upload: pad kernel: %s", hipGetErrorString(e));
  }
  DevBrick b{};
  for (int a = 0; a < 3; ++a) {
    b.space_lo[a] = desc->space_range.lower[a];
    b.space_hi[a] = desc->space_range.upper[a];
    b.vox_lo[a] = (float)desc->voxel_range.lower[a];
    b.n[a] = n[a];
  }
  b.px = (int32_t)px;
  b.pxy = (int32_t)(px * py);
  b.dtype = store;
  b.origin = mem;  // padded element (-1,-1,-1)
  ctx->bricks.push_back(b);
  ctx->brick_mem.push_back(mem);
  ctx->descs.push_back(*desc);
  ctx->brick_bytes += mem_bytes;
  ctx->bricks_dirty = true;
  // the device brick table follows at once, so render calls find it current
  // (and aloha_render_frame_noalloc never has to grow it); if it cannot,
  // the brick is not added
  if (int rc = sync_bricks(ctx)) {
    ctx->bricks.pop_back();
    ctx->brick_mem.pop_back();
    ctx->descs.pop_back();
    ctx->brick_bytes -= mem_bytes;
    hipFree(mem);
    return rc;
  }
  if (brick_id) *brick_id = (int32_t)ctx->bricks.size() - 1;
  return ALOHA_OK;
}

extern "C" int aloha_brick_count(const aloha_ctx *ctx, int32_t *count) {
  if (!ctx || !count) return fail(nullptr, ALOHA_ERR_ARG, "aloha_brick_count: NULL");
  *count = (int32_t)ctx->bricks.size();
  return ALOHA_OK;
}

extern "C" int aloha_brick_bytes(const aloha_ctx *ctx, uint64_t *bytes) {
  if (!ctx || !bytes) return fail(nullptr, ALOHA_ERR_ARG, "aloha_brick_bytes: NULL");
  *bytes = ctx->brick_bytes;
  return ALOHA_OK;
}

extern "C" int aloha_set_u8_storage(aloha_ctx *ctx, int32_t keep_u8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_set_u8_storage: ctx is NULL");
  ctx->u8_storage = keep_u8 != 0;
  return ALOHA_OK;
}

extern "C" int aloha_brick_dtype(const aloha_ctx *ctx, int32_t brick_id, int32_t *dtype) {
  if (!ctx || !dtype) return fail(nullptr, ALOHA_ERR_ARG, "aloha_brick_dtype: NULL");
  if (brick_id < 0 || brick_id >= (int)ctx->bricks.size())
    return fail(const_cast<aloha_ctx *>(ctx), ALOHA_ERR_ARG, "aloha_brick_dtype: bad brick id %d",
                brick_id);
  *dtype = ctx->bricks[brick_id].dtype;
  return ALOHA_OK;
}

extern "C" int aloha_tf_set(aloha_ctx *ctx, const float *rgba, int32_t n) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_tf_set: ctx is NULL");
  if (!rgba || n < 1 || n > ALOHA_MAX_TF)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_tf_set: need 1..%d entries", ALOHA_MAX_TF);
  for (int i = 0; i < 4 * n; ++i)
    if (!std::isfinite(rgba[i]))
      return fail(ctx, ALOHA_ERR_ARG, "aloha_tf_set: entry %d is not finite", i / 4);
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // ordered after earlier launches on the stream that read the old table
  HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (n > ctx->tf_cap) {
    if (ctx->d_tf) HIP_CHECK(ctx, hipFree(ctx->d_tf));
    ctx->d_tf = nullptr;
    ctx->tf_cap = ctx->tf_n = 0;  // no table until the allocation succeeds
    ++ctx->gen;
    HIP_CHECK(ctx, dev_malloc((void **)&ctx->d_tf, sizeof(float4) * (size_t)n));
    ctx->tf_cap = n;
    ++ctx->gen;
  }
  HIP_CHECK(ctx, hipMemcpy(ctx->d_tf, rgba, sizeof(float4) * (size_t)n, hipMemcpyHostToDevice));
  if (n != ctx->tf_n) ++ctx->gen;
  ctx->tf_n = n;
  return ALOHA_OK;
}

extern "C" int aloha_camera_set(aloha_ctx *ctx, const float org[3], const float dir_00[3],
                                const float dir_du[3], const float dir_dv[3]) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_camera_set: ctx is NULL");
  if (!org || !dir_00 || !dir_du || !dir_dv) return fail(ctx, ALOHA_ERR_ARG, "aloha_camera_set: NULL");
  if (!finite3(org) || !finite3(dir_00) || !finite3(dir_du) || !finite3(dir_dv))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_camera_set: non-finite camera");
  std::memcpy(ctx->cam, org, 12);
  std::memcpy(ctx->cam + 3, dir_00, 12);
  std::memcpy(ctx->cam + 6, dir_du, 12);
  std::memcpy(ctx->cam + 9, dir_dv, 12);
  ctx->has_cam = true;
  return ALOHA_OK;
}

namespace aloha {

int prepare_ranks(aloha_ctx *ctx, const MarchArgs &a, int lanes, RanksArgs &q, int &waves) {
  int64_t total = 0;
  q.run_off[0] = 0;
  for (int s = 0; s < a.P; ++s) {
    const int32_t *rc = a.rect[s];
    if (rc[0] < rc[2] && rc[1] < rc[3])
      total += (int64_t)(rc[3] - rc[1]) * ((rc[2] - rc[0] + kRankRun - 1) / kRankRun);
    if (total > INT32_MAX / 128)
      return fail(ctx, ALOHA_ERR_LIMIT, "render: rank rectangles too large");
    q.run_off[s + 1] = (int32_t)total;
  }
  q.frame = nullptr;
  q.frame8 = nullptr;
  q.box = make_int4(0, 0, 0, 0);
  q.compact = 0;
  // persistent grid: as many one-wave blocks as the GPU holds at once
  const int li = (lanes == 1 ? 0 : lanes == 2 ? 1 : lanes == 4 ? 2 : 3) + (a.u8_lut ? 4 : 0);
  const size_t lds = ranks_lds_bytes(a);
  if (ctx->q_waves[li] == 0 || ctx->q_waves_lds[li] != lds) {
    int per_cu = 0, cus = 0;
    HIP_CHECK(ctx, ranks_occupancy(lanes, a.u8_lut != 0, lds, &per_cu));
    HIP_CHECK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->q_waves[li] = std::max(1, per_cu) * std::max(1, cus);
    ctx->q_waves_lds[li] = lds;
  }
  waves = ctx->q_waves[li];
  return ALOHA_OK;
}

}  // namespace aloha

extern "C" int aloha_render_ranks(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                  int32_t H, int32_t P, const int32_t *rank_offsets,
                                  const int32_t *brick_ids, const aloha_images *out) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_ranks: ctx is NULL");
  if (!out) return fail(ctx, ALOHA_ERR_ARG, "aloha_render_ranks: out is NULL");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  if (params->accum_id > 0 && !out->accum)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_ranks: accum_id > 0 needs the accum buffer");
  a.rgba = (float4 *)out->rgba;
  a.rgba8 = out->rgba8;
  a.depth = out->depth;
  a.accum = (float4 *)out->accum;
  a.nsamples = out->nsamples;
  RanksArgs q;
  int waves = 0;
  if (int rc = prepare_ranks(ctx, a, ctx->march_lanes, q, waves)) return rc;
  HIP_CHECK(ctx, launch_march_ranks(a, q, ctx->march_lanes, waves, ctx->stream));
  return ALOHA_OK;
}

namespace {
// validates order[] as a permutation of the P slots and copies it into c (defined below)
int check_order(aloha_ctx *ctx, const int32_t *order, int32_t P, CompositeArgs &c);
}  // namespace

// Renderer::render (waikikiRenderer.cpp:1015-1083): the raygen launch for
// every rank, then IceT's ordered blend of their images.  For a plain frame
// the per-rank launch also writes the composite outputs transparent outside
// the rectangles' bounding box (in its fill phase, which overlaps the march
// of other waves), and the sparse composite then covers only that box.  The
// results equal aloha_render_ranks followed by aloha_composite_over_sparse
// over the rectangles of aloha_rank_rects, bit for bit.
extern "C" int aloha_render_frame(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                  int32_t H, int32_t P, const int32_t *rank_offsets,
                                  const int32_t *brick_ids, const aloha_images *images,
                                  const int32_t *order, float *d_out_rgba,
                                  uint32_t *d_out_rgba8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_frame: ctx is NULL");
  if (!images)  // no rank images wanted: compact ones in the context's workspace
    return frame_compact(ctx, "aloha_render_frame", false, params, W, H, P, rank_offsets,
                         brick_ids, order, d_out_rgba, d_out_rgba8);
  if (!images->rgba)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_frame: the rank images need float RGBA");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  if (params->accum_id > 0 && !images->accum)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_frame: accum_id > 0 needs the accum buffer");
  CompositeArgs c;
  if (int rc = check_order(ctx, order, P, c)) return rc;
  c.npix = (int64_t)W * H;
  a.rgba = (float4 *)images->rgba;
  a.rgba8 = images->rgba8;
  a.depth = images->depth;
  a.accum = (float4 *)images->accum;
  a.nsamples = images->nsamples;
  RanksArgs q;
  int waves = 0;
  if (int rc = prepare_ranks(ctx, a, ctx->march_lanes, q, waves)) return rc;
  // as aloha_rank_rects: no pixel is known transparent with a heat map or
  // accumulation (then the composite covers the whole image)
  const bool whole = (params->flags & ALOHA_FLAG_HEATMAP) || params->accum_id > 0;
  int x0 = W, y0 = H, x1 = 0, y1 = 0;  // bounding box of the rectangles
  for (int s = 0; s < P; ++s) {
    const int32_t full[4] = {0, 0, W, H};
    for (int k = 0; k < 4; ++k) c.rect[s][k] = whole ? full[k] : a.rect[s][k];
    const int32_t *rc = c.rect[s];
    if (rc[0] >= rc[2] || rc[1] >= rc[3]) continue;
    x0 = std::min(x0, (int)rc[0]), y0 = std::min(y0, (int)rc[1]);
    x1 = std::max(x1, (int)rc[2]), y1 = std::max(y1, (int)rc[3]);
  }
  if (x0 >= x1 || y0 >= y1) x0 = y0 = x1 = y1 = 0;  // no rectangle: no composite launch
  q.frame = (float4 *)d_out_rgba;
  q.frame8 = d_out_rgba8;
  q.box = make_int4(x0, y0, x1, y1);
  HIP_CHECK(ctx, launch_march_ranks(a, q, ctx->march_lanes, waves, ctx->stream));
  HIP_CHECK(ctx, launch_composite_over_sparse((const float4 *)images->rgba, c, W, x0, y0, x1, y1,
                                              (float4 *)d_out_rgba, d_out_rgba8, ctx->stream));
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free frame: the rank images live in a workspace the library
// holds, not in caller memory.
// ---------------------------------------------------------------------------
// The ranks' compact float RGBA images of aloha_render_frame_noalloc.  Each
// GPU holds its own copy for the life of the process.  In host code the name
// is not a device address: the address of the current device's copy comes
// from hipGetSymbolAddress (scratch_on_device) and only that is used.
static __device__ __align__(256) float4 scratch[ALOHA_NOALLOC_SCRATCH_BYTES / sizeof(float4)];

namespace aloha {
namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

// Address of `device`'s scratch array; looked up once per device (with that
// device current: the caller has made it so).
int scratch_on_device(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace aloha

namespace aloha {
// The frame with compact rank images: the per-rank launch stores each rank's
// pixels only inside its rectangle, the rectangles one after the other
// (row-major each: IceT's sparse images, kept on the device), then the
// composite of the rectangles' bounding box.  noalloc: the workspace is the
// static per-device array (ALOHA_NOALLOC_SCRATCH_BYTES, nothing allocated,
// the brick table must be current); else the context's own workspace, grown
// when a frame needs more.
int frame_compact(aloha_ctx *ctx, const char *name, bool noalloc,
                  const aloha_render_params *params, int32_t W, int32_t H, int32_t P,
                  const int32_t *rank_offsets, const int32_t *brick_ids, const int32_t *order,
                  float *d_out_rgba, uint32_t *d_out_rgba8) {
  if (noalloc && !ctx->scratch)
    return fail(ctx, ALOHA_ERR_STATE, "%s: no workspace on device %d (%s)", name, ctx->device,
                ctx->scratch_err.c_str());
  if (!params) return fail(ctx, ALOHA_ERR_ARG, "%s: params is NULL", name);
  if (params->accum_id != 0 || (params->flags & ALOHA_FLAG_HEATMAP))
    return fail(ctx, ALOHA_ERR_ARG, "%s: accum_id 0 and no heat map only (no accum buffer)", name);
  // build_march would grow a stale device brick table; uploads keep it current
  if (noalloc && ctx->bricks_dirty)
    return fail(ctx, ALOHA_ERR_STATE, "%s: brick table not on the device", name);
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  CompositeArgs c;
  if (int rc = check_order(ctx, order, P, c)) return rc;
  c.npix = (int64_t)W * H;
  // compact layout: the rectangles one after the other, each row-major
  int64_t total = 0;
  int x0 = W, y0 = H, x1 = 0, y1 = 0;  // bounding box of the rectangles
  for (int s = 0; s < P; ++s) {
    const int32_t *rc = a.rect[s];
    for (int k = 0; k < 4; ++k) c.rect[s][k] = rc[k];
    c.cbase[s] = (int32_t)std::min<int64_t>(total, INT32_MAX);
    if (rc[0] >= rc[2] || rc[1] >= rc[3]) continue;
    total += (int64_t)(rc[2] - rc[0]) * (rc[3] - rc[1]);
    x0 = std::min(x0, (int)rc[0]), y0 = std::min(y0, (int)rc[1]);
    x1 = std::max(x1, (int)rc[2]), y1 = std::max(y1, (int)rc[3]);
  }
  if (x0 >= x1 || y0 >= y1) x0 = y0 = x1 = y1 = 0;  // no rectangle: no composite launch
  if (total > INT32_MAX) return fail(ctx, ALOHA_ERR_LIMIT, "%s: rectangles too large", name);
  float4 *ws = nullptr;
  if (noalloc) {
    if (total * (int64_t)sizeof(float4) > (int64_t)ALOHA_NOALLOC_SCRATCH_BYTES)
      return fail(ctx, ALOHA_ERR_LIMIT,
                  "%s: the ranks' rectangles hold %lld pixels (%lld bytes); the workspace holds "
                  "%lld bytes",
                  name, (long long)total, (long long)(total * (int64_t)sizeof(float4)),
                  (long long)ALOHA_NOALLOC_SCRATCH_BYTES);
    ws = (float4 *)ctx->scratch;
  } else {
    if ((size_t)total > ctx->frame_ws_cap) {
      HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // earlier frames read the old one
      if (ctx->frame_ws) HIP_CHECK(ctx, hipFree(ctx->frame_ws));
      ctx->frame_ws = nullptr;
      ctx->frame_ws_cap = 0;
      const size_t cap = (size_t)total + (size_t)total / 4 + 1024;
      HIP_CHECK(ctx, dev_malloc((void **)&ctx->frame_ws, cap * sizeof(float4)));
      ctx->frame_ws_cap = cap;
    }
    ws = ctx->frame_ws;
  }
  a.rgba = ws;
  a.rgba8 = nullptr;
  a.depth = nullptr;
  a.accum = nullptr;
  a.nsamples = nullptr;
  RanksArgs q;
  int waves = 0;
  if (int rc = prepare_ranks(ctx, a, ctx->march_lanes, q, waves)) return rc;
  q.compact = 1;
  for (int s = 0; s < P; ++s) q.cbase[s] = c.cbase[s];
  q.frame = (float4 *)d_out_rgba;
  q.frame8 = d_out_rgba8;
  q.box = make_int4(x0, y0, x1, y1);
  HIP_CHECK(ctx, launch_march_ranks(a, q, ctx->march_lanes, waves, ctx->stream));
  HIP_CHECK(ctx, launch_composite_over_compact(ws, c, W, x0, y0, x1, y1, (float4 *)d_out_rgba,
                                               d_out_rgba8, ctx->stream));
  return ALOHA_OK;
}
}  // namespace aloha

extern "C" int aloha_render_frame_noalloc(aloha_ctx *ctx, const aloha_render_params *params,
                                          int32_t W, int32_t H, int32_t P,
                                          const int32_t *rank_offsets, const int32_t *brick_ids,
                                          const int32_t *order, float *d_out_rgba,
                                          uint32_t *d_out_rgba8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_frame_noalloc: ctx is NULL");
  return frame_compact(ctx, "aloha_render_frame_noalloc", true, params, W, H, P, rank_offsets,
                       brick_ids, order, d_out_rgba, d_out_rgba8);
}

extern "C" int aloha_render_composite(aloha_ctx *ctx, const aloha_render_params *params,
                                      int32_t W, int32_t H, int32_t P,
                                      const int32_t *rank_offsets, const int32_t *brick_ids,
                                      const int32_t *order, float *d_out_rgba,
                                      uint32_t *d_out_rgba8, uint32_t *d_nsamples) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_composite: ctx is NULL");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  if (params->accum_id != 0 || (params->flags & (ALOHA_FLAG_HEATMAP | ALOHA_FLAG_CROSSHAIR)))
    return fail(ctx, ALOHA_ERR_ARG,
                "aloha_render_composite: accum_id 0 and no heat map / crosshair only");
  if (!order) return fail(ctx, ALOHA_ERR_ARG, "aloha_render_composite: order is NULL");
  std::vector<char> seen(P, 0);
  for (int k = 0; k < P; ++k) {
    if (order[k] < 0 || order[k] >= P || seen[order[k]])
      return fail(ctx, ALOHA_ERR_ARG, "aloha_render_composite: order is not a permutation");
    seen[order[k]] = 1;
    a.order[k] = order[k];
  }
  a.rgba = (float4 *)d_out_rgba;
  a.rgba8 = d_out_rgba8;
  a.nsamples = d_nsamples;
  HIP_CHECK(ctx, launch_march_composite(a, ctx->fused_lanes, ctx->stream));
  return ALOHA_OK;
}

namespace {
bool lanes_ok(int n) { return n == 1 || n == 2 || n == 4 || n == 8; }
}  // namespace

extern "C" int aloha_set_march_schedule(aloha_ctx *ctx, int32_t ranks_lanes, int32_t fused_lanes) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_set_march_schedule: ctx is NULL");
  if (!lanes_ok(ranks_lanes) || !lanes_ok(fused_lanes))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_set_march_schedule: lanes per ray must be 1, 2, 4 or 8");
  ctx->march_lanes = ranks_lanes;
  ctx->fused_lanes = fused_lanes;
  return ALOHA_OK;
}

extern "C" int aloha_get_march_schedule(const aloha_ctx *ctx, int32_t *ranks_lanes,
                                        int32_t *fused_lanes) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_get_march_schedule: ctx is NULL");
  if (ranks_lanes) *ranks_lanes = ctx->march_lanes;
  if (fused_lanes) *fused_lanes = ctx->fused_lanes;
  return ALOHA_OK;
}

extern "C" int aloha_autotune_march(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                    int32_t H, int32_t P, const int32_t *rank_offsets,
                                    const int32_t *brick_ids, const int32_t *order,
                                    const aloha_images *images, float *d_out_rgba,
                                    uint32_t *d_out_rgba8, int32_t reps, float *table,
                                    int32_t *rows) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_autotune_march: ctx is NULL");
  if (!images || reps < 1) return fail(ctx, ALOHA_ERR_ARG, "aloha_autotune_march: bad argument");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  if (params->accum_id > 0 && !images->accum)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_autotune_march: accum_id > 0 needs the accum buffer");
  const bool fused = order && (d_out_rgba || d_out_rgba8);
  MarchArgs f = a;
  if (fused) {
    if (params->accum_id != 0 || (params->flags & (ALOHA_FLAG_HEATMAP | ALOHA_FLAG_CROSSHAIR)))
      return fail(ctx, ALOHA_ERR_ARG, "aloha_autotune_march: the fused form needs accum_id 0, no flags");
    std::vector<char> seen(P, 0);
    for (int k = 0; k < P; ++k) {
      if (order[k] < 0 || order[k] >= P || seen[order[k]])
        return fail(ctx, ALOHA_ERR_ARG, "aloha_autotune_march: order is not a permutation");
      seen[order[k]] = 1;
      f.order[k] = order[k];
    }
    f.rgba = (float4 *)d_out_rgba;
    f.rgba8 = d_out_rgba8;
  }
  a.rgba = (float4 *)images->rgba;
  a.rgba8 = images->rgba8;
  a.depth = images->depth;
  a.accum = (float4 *)images->accum;
  a.nsamples = images->nsamples;
  hipEvent_t e0, e1;
  HIP_CHECK(ctx, hipEventCreate(&e0));
  HIP_CHECK(ctx, hipEventCreate(&e1));
  auto time = [&](auto &&launch, float &ms) {
    hipError_t e = launch();  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    for (int r = 0; e == hipSuccess && r < reps; ++r) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    return e;
  };
  const int opts[4] = {1, 2, 4, 8};
  float best_r = 1e30f, best_f = 1e30f;
  int lr = ctx->march_lanes, lf = ctx->fused_lanes;
  hipError_t e = hipSuccess;
  for (int i = 0; i < 4 && e == hipSuccess; ++i) {
    const int g = opts[i];
    float mr = 0.f, mf = 0.f;
    RanksArgs q;
    int waves = 0;
    if (int rc = prepare_ranks(ctx, a, g, q, waves)) return rc;
    e = time([&] { return launch_march_ranks(a, q, g, waves, ctx->stream); }, mr);
    if (e == hipSuccess && fused) e = time([&] { return launch_march_composite(f, g, ctx->stream); }, mf);
    if (e != hipSuccess) break;
    if (table) {
      table[3 * i] = (float)g;
      table[3 * i + 1] = mr;
      table[3 * i + 2] = fused ? mf : 0.f;
    }
    if (mr < best_r) best_r = mr, lr = g;
    if (fused && mf < best_f) best_f = mf, lf = g;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(ctx, ALOHA_ERR_HIP, "aloha_autotune_march: %s", hipGetErrorString(e));
  ctx->march_lanes = lr;
  ctx->fused_lanes = lf;
  if (rows) *rows = 4;
  return ALOHA_OK;
}

namespace {
// The footprint pass: per brick id, the distinct voxels the march of the
// launch reads (bitset per brick, counted on the host).
int footprint(aloha_ctx *ctx, const char *name, const aloha_render_params *params, int32_t W,
              int32_t H, int32_t P, const int32_t *rank_offsets, const int32_t *brick_ids,
              std::vector<uint64_t> &per_brick) {
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  // one bitset per brick, word-aligned
  uint64_t words = 0;
  for (auto &b : ctx->bricks) {
    const uint64_t nv = (uint64_t)b.n[0] * b.n[1] * b.n[2];
    if (words + (nv + 31) / 32 > 0x7fffffffull)
      return fail(ctx, ALOHA_ERR_LIMIT, "%s: bitset too large", name);
    b.touched_word = (int32_t)words;
    words += (nv + 31) / 32;
  }
  ctx->bricks_dirty = true;
  if (int rc = sync_bricks(ctx)) return rc;
  uint32_t *bits = nullptr;
  HIP_CHECK(ctx, dev_malloc((void **)&bits, words * 4));
  hipError_t e = hipMemsetAsync(bits, 0, words * 4, ctx->stream);
  a.touched = bits;
  a.bricks = ctx->d_bricks;
  if (e == hipSuccess) e = launch_march_footprint(a, ctx->stream);
  std::vector<uint32_t> host(words);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(host.data(), bits, words * 4, hipMemcpyDeviceToHost);
  hipFree(bits);
  if (e != hipSuccess) return fail(ctx, ALOHA_ERR_HIP, "%s: %s", name, hipGetErrorString(e));
  per_brick.assign(ctx->bricks.size(), 0);
  for (size_t k = 0; k < ctx->bricks.size(); ++k) {
    const auto &b = ctx->bricks[k];
    const uint64_t w0 = (uint64_t)b.touched_word;
    const uint64_t nw = ((uint64_t)b.n[0] * b.n[1] * b.n[2] + 31) / 32;
    for (uint64_t w = w0; w < w0 + nw; ++w) per_brick[k] += (uint64_t)__builtin_popcount(host[w]);
  }
  return ALOHA_OK;
}
}  // namespace

extern "C" int aloha_render_footprint(aloha_ctx *ctx, const aloha_render_params *params,
                                      int32_t W, int32_t H, int32_t P,
                                      const int32_t *rank_offsets, const int32_t *brick_ids,
                                      uint64_t *touched_voxels) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_footprint: ctx is NULL");
  if (!touched_voxels) return fail(ctx, ALOHA_ERR_ARG, "aloha_render_footprint: NULL output");
  std::vector<uint64_t> per;
  if (int rc = footprint(ctx, "aloha_render_footprint", params, W, H, P, rank_offsets, brick_ids,
                         per))
    return rc;
  uint64_t cnt = 0;
  for (uint64_t v : per) cnt += v;
  *touched_voxels = cnt;
  return ALOHA_OK;
}

extern "C" int aloha_render_footprint_bricks(aloha_ctx *ctx, const aloha_render_params *params,
                                             int32_t W, int32_t H, int32_t P,
                                             const int32_t *rank_offsets,
                                             const int32_t *brick_ids, uint64_t *touched,
                                             int32_t cap) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_footprint_bricks: ctx is NULL");
  if (!touched || cap < (int32_t)ctx->bricks.size())
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_footprint_bricks: need %d outputs",
                (int)ctx->bricks.size());
  std::vector<uint64_t> per;
  if (int rc = footprint(ctx, "aloha_render_footprint_bricks", params, W, H, P, rank_offsets,
                         brick_ids, per))
    return rc;
  for (size_t k = 0; k < per.size(); ++k) touched[k] = per[k];
  return ALOHA_OK;
}

namespace {
int check_order(aloha_ctx *ctx, const int32_t *order, int32_t P, CompositeArgs &c) {
  if (P < 1 || P > ALOHA_MAX_RANKS)
    return fail(ctx, ALOHA_ERR_LIMIT, "composite: P=%d outside [1,%d]", P, ALOHA_MAX_RANKS);
  if (!order) return fail(ctx, ALOHA_ERR_ARG, "composite: order is NULL");
  std::vector<char> seen(P, 0);
  for (int k = 0; k < P; ++k) {
    if (order[k] < 0 || order[k] >= P || seen[order[k]])
      return fail(ctx, ALOHA_ERR_ARG, "composite: order is not a permutation of 0..%d", P - 1);
    seen[order[k]] = 1;
    c.order[k] = order[k];
  }
  c.P = P;
  return ALOHA_OK;
}
}  // namespace

extern "C" int aloha_composite_over(aloha_ctx *ctx, const float *d_rank_rgba, int32_t P,
                                    const int32_t *order, int32_t W, int32_t H,
                                    float *d_out_rgba, uint32_t *d_out_rgba8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_composite_over: ctx is NULL");
  if (!d_rank_rgba || W < 1 || H < 1) return fail(ctx, ALOHA_ERR_ARG, "aloha_composite_over: bad argument");
  CompositeArgs c;
  if (int rc = check_order(ctx, order, P, c)) return rc;
  c.npix = (int64_t)W * H;
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  HIP_CHECK(ctx, launch_composite_over((const float4 *)d_rank_rgba, c, (float4 *)d_out_rgba,
                                       d_out_rgba8, ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_composite_over_sparse(aloha_ctx *ctx, const float *d_rank_rgba, int32_t P,
                                           const int32_t *order, const int32_t *rects, int32_t W,
                                           int32_t H, float *d_out_rgba, uint32_t *d_out_rgba8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_composite_over_sparse: ctx is NULL");
  if (!d_rank_rgba || !rects || W < 1 || H < 1)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_composite_over_sparse: bad argument");
  CompositeArgs c;
  if (int rc = check_order(ctx, order, P, c)) return rc;
  c.npix = (int64_t)W * H;
  for (int s = 0; s < P; ++s)
    for (int k = 0; k < 4; ++k) c.rect[s][k] = rects[4 * s + k];
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  HIP_CHECK(ctx, launch_composite_over_sparse((const float4 *)d_rank_rgba, c, W, 0, 0, W, H,
                                              (float4 *)d_out_rgba, d_out_rgba8, ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_rank_rects(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                int32_t H, int32_t P, const int32_t *rank_offsets,
                                const int32_t *brick_ids, int32_t *rects) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_rank_rects: ctx is NULL");
  if (!rects) return fail(ctx, ALOHA_ERR_ARG, "aloha_rank_rects: rects is NULL");
  MarchArgs a;
  if (int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, a)) return rc;
  // a heat map colours every pixel, and accumulation adds earlier frames'
  // pixels: then no pixel is known to be transparent
  const bool whole = (params->flags & ALOHA_FLAG_HEATMAP) || params->accum_id > 0;
  for (int s = 0; s < P; ++s) {
    const int32_t full[4] = {0, 0, W, H};
    for (int k = 0; k < 4; ++k) rects[4 * s + k] = whole ? full[k] : a.rect[s][k];
  }
  return ALOHA_OK;
}

extern "C" int aloha_composite_over_rgba8(aloha_ctx *ctx, const uint32_t *d_rank_rgba8,
                                          int32_t P, const int32_t *order, int32_t W, int32_t H,
                                          uint32_t *d_out_rgba8) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_composite_over_rgba8: ctx is NULL");
  if (!d_rank_rgba8 || !d_out_rgba8 || W < 1 || H < 1)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_composite_over_rgba8: bad argument");
  CompositeArgs c;
  if (int rc = check_order(ctx, order, P, c)) return rc;
  c.npix = (int64_t)W * H;
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  HIP_CHECK(ctx, launch_composite_over_rgba8(d_rank_rgba8, c, d_out_rgba8, ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_composite_zmin(aloha_ctx *ctx, const uint32_t *d_rank_rgba8,
                                    const float *d_rank_depth, int32_t P, int32_t W, int32_t H,
                                    uint32_t *d_out_rgba8, float *d_out_depth) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_composite_zmin: ctx is NULL");
  if (!d_rank_rgba8 || !d_rank_depth || W < 1 || H < 1 || P < 1)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_composite_zmin: bad argument");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  HIP_CHECK(ctx, launch_composite_zmin(d_rank_rgba8, d_rank_depth, P, (int64_t)W * H,
                                       d_out_rgba8, d_out_depth, ctx->stream));
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// device-resident inputs
// ---------------------------------------------------------------------------
extern "C" int aloha_tf_set_device(aloha_ctx *ctx, const float *d_rgba, int32_t n) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_tf_set_device: ctx is NULL");
  if (!d_rgba || n < 1 || n > ALOHA_MAX_TF)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_tf_set_device: need 1..%d entries", ALOHA_MAX_TF);
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (n > ctx->tf_cap) {
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_tf) HIP_CHECK(ctx, hipFree(ctx->d_tf));
    ctx->d_tf = nullptr;
    ctx->tf_cap = ctx->tf_n = 0;  // no table until the allocation succeeds
    ++ctx->gen;
    HIP_CHECK(ctx, dev_malloc((void **)&ctx->d_tf, sizeof(float4) * (size_t)n));
    ctx->tf_cap = n;
    ++ctx->gen;
  }
  HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_tf, d_rgba, sizeof(float4) * (size_t)n,
                                hipMemcpyDeviceToDevice, ctx->stream));
  if (n != ctx->tf_n) ++ctx->gen;
  ctx->tf_n = n;
  return ALOHA_OK;
}

extern "C" int aloha_brick_update(aloha_ctx *ctx, int32_t brick_id, const void *d_voxels,
                                  int32_t dtype) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_brick_update: ctx is NULL");
  if (brick_id < 0 || brick_id >= (int)ctx->bricks.size() || !d_voxels)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_update: bad brick id %d or NULL voxels", brick_id);
  if (dtype != ALOHA_U8 && dtype != ALOHA_U16 && dtype != ALOHA_F32)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_brick_update: unknown dtype %d", dtype);
  const DevBrick &b = ctx->bricks[brick_id];
  if (b.dtype == ALOHA_U8 && dtype != ALOHA_U8)
    return fail(ctx, ALOHA_ERR_ARG,
                "aloha_brick_update: brick %d keeps u8 voxels; its update must be u8 too", brick_id);
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  HIP_CHECK(ctx, launch_pad_brick(d_voxels, dtype, b.n, ctx->brick_mem[brick_id], b.dtype, b.px,
                                  b.pxy / b.px, b.n[2] + kPadLo + kPadHi, ctx->stream));
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// differentiable form
// ---------------------------------------------------------------------------
namespace {
int build_diff(aloha_ctx *ctx, const aloha_render_params *p, int32_t W, int32_t H, int32_t V,
               const float *cams, int32_t nb, const int32_t *brick_ids, DiffArgs &a) {
  if (!p) return fail(ctx, ALOHA_ERR_ARG, "diff: params is NULL");
  if (!(p->dt > 0.f) || !std::isfinite(p->dt))
    return fail(ctx, ALOHA_ERR_ARG, "diff: dt must be finite and > 0");
  if (p->jitter_mode != ALOHA_JITTER_RANDOM && p->jitter_mode != ALOHA_JITTER_FIXED)
    return fail(ctx, ALOHA_ERR_ARG, "diff: unknown jitter_mode %d", p->jitter_mode);
  if (p->jitter_mode == ALOHA_JITTER_FIXED && !(p->jitter_value >= 0.f && p->jitter_value < 1.f))
    return fail(ctx, ALOHA_ERR_ARG, "diff: jitter_value must be in [0,1)");
  if (W < 1 || H < 1 || (int64_t)W * H * V > (int64_t)1 << 30)
    return fail(ctx, ALOHA_ERR_ARG, "diff: image %dx%dx%d out of range", W, H, V);
  if (V < 1 || V > kMaxViews || !cams)
    return fail(ctx, ALOHA_ERR_LIMIT, "diff: V=%d outside [1,%d]", V, kMaxViews);
  if (nb < 1 || nb > kMaxDiffBricks || !brick_ids)
    return fail(ctx, ALOHA_ERR_LIMIT, "diff: nb=%d outside [1,%d]", nb, kMaxDiffBricks);
  for (int k = 0; k < nb; ++k) {
    if (brick_ids[k] < 0 || brick_ids[k] >= (int)ctx->bricks.size())
      return fail(ctx, ALOHA_ERR_ARG, "diff: brick id %d not uploaded", brick_ids[k]);
    if (ctx->bricks[brick_ids[k]].dtype != ALOHA_F32)
      return fail(ctx, ALOHA_ERR_ARG,
                  "diff: brick %d keeps u8 voxels; the differentiable form needs float32 bricks "
                  "(upload it after aloha_set_u8_storage(ctx, 0))", brick_ids[k]);
  }
  for (int i = 0; i < 12 * V; ++i)
    if (!std::isfinite(cams[i])) return fail(ctx, ALOHA_ERR_ARG, "diff: non-finite camera");
  for (int v = 0; v < V; ++v)
    if (int rc = check_steps(ctx, cams + 12 * v, ctx->bricks, brick_ids, nb, p->dt)) return rc;
  if (!ctx->d_tf || ctx->tf_n < 1) return fail(ctx, ALOHA_ERR_STATE, "diff: no transfer function set");
  if (int rc = sync_bricks(ctx)) return rc;
  std::memset(&a, 0, sizeof a);
  std::memcpy(a.cam, cams, sizeof(float) * 12 * (size_t)V);
  a.dt = p->dt;
  a.jitter_value = p->jitter_value;
  a.jitter_mode = p->jitter_mode;
  a.alpha_max = 0.9999f;
  a.W = W;
  a.H = H;
  a.V = V;
  a.tf_n = ctx->tf_n;
  a.nb = nb;
  for (int k = 0; k < nb; ++k) a.brick_ids[k] = brick_ids[k];
  a.bricks = ctx->d_bricks;
  a.tf = ctx->d_tf;
  // the forward's 16x16 blocks: those of each view's brick rectangle
  a.block_off[0] = 0;
  for (int v = 0; v < V; ++v) {
    int32_t r[4];
    rank_rect(cams + 12 * v, ctx->bricks, brick_ids, nb, W, H, r);
    int32_t *fb = a.fbox[v];
    fb[0] = r[0] / 16;
    fb[1] = r[1] / 16;
    fb[2] = r[2] > r[0] ? (r[2] + 15) / 16 - fb[0] : 0;
    fb[3] = r[3] > r[1] ? (r[3] + 15) / 16 - fb[1] : 0;
    if (fb[2] == 0 || fb[3] == 0) fb[2] = fb[3] = 0;
    a.block_off[v + 1] = a.block_off[v] + fb[2] * fb[3];
  }
  return ALOHA_OK;
}
}  // namespace

namespace {
// Bins of the brick slots: 16^3 blocks over the corners -1 .. n of each axis
// (every slot, with or without a gradient grid: the counts cover them all).
// Returns the bin total; max_dims / max_voxels over the slots with a grid.
int64_t diff_layout(const aloha_ctx *ctx, int32_t nb, const int32_t *brick_ids,
                    float *const *dvox, DiffBins &b, int max_dims[3], int64_t &max_voxels) {
  std::memset(&b, 0, sizeof b);
  int64_t nbins = 0;
  max_voxels = 0;
  max_dims[0] = max_dims[1] = max_dims[2] = 0;
  for (int k = 0; k < nb; ++k) {
    b.base[k] = (int32_t)std::min<int64_t>(nbins, INT32_MAX);
    const DevBrick &db = ctx->bricks[brick_ids[k]];
    b.nbx[k] = (db.n[0] + 1) / kBinEdge + 1;
    b.nby[k] = (db.n[1] + 1) / kBinEdge + 1;
    b.nbz[k] = (db.n[2] + 1) / kBinEdge + 1;
    nbins += (int64_t)b.nbx[k] * b.nby[k] * b.nbz[k];
    if (dvox && dvox[k]) {
      max_voxels = std::max<int64_t>(max_voxels, (int64_t)db.n[0] * db.n[1] * db.n[2]);
      for (int ax = 0; ax < 3; ++ax) max_dims[ax] = std::max(max_dims[ax], db.n[ax]);
    }
  }
  b.base[nb] = (int32_t)std::min<int64_t>(nbins, INT32_MAX);
  b.nbins = (int32_t)std::min<int64_t>(nbins, INT32_MAX);
  return nbins;
}

constexpr int64_t kMaxDiffBins = 1 << 24;
}  // namespace

extern "C" int aloha_diff_forward(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                  int32_t H, int32_t V, const float *cams, int32_t nb,
                                  const int32_t *brick_ids, float *d_rgba,
                                  uint32_t *d_nsamples) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_diff_forward: ctx is NULL");
  if (!d_rgba || !d_nsamples) return fail(ctx, ALOHA_ERR_ARG, "aloha_diff_forward: NULL output");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DiffArgs a;
  if (int rc = build_diff(ctx, params, W, H, V, cams, nb, brick_ids, a)) return rc;
  a.rgba = (float4 *)d_rgba;
  a.nsamples = d_nsamples;
  HIP_CHECK(ctx, launch_diff(a, false, ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_diff_backward(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                   int32_t H, int32_t V, const float *cams, int32_t nb,
                                   const int32_t *brick_ids, const float *d_rgba,
                                   const uint32_t *d_nsamples, const float *d_dl_drgba,
                                   float *d_dtf, float *const *d_dvoxels) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_diff_backward: ctx is NULL");
  if (!d_rgba || !d_nsamples || !d_dl_drgba)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_diff_backward: NULL input");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DiffArgs a;
  if (int rc = build_diff(ctx, params, W, H, V, cams, nb, brick_ids, a)) return rc;
  if ((size_t)sizeof(float4) * (a.tf_n + 4) + sizeof(float) * 4 * (size_t)a.tf_n > 64 * 1024)
    return fail(ctx, ALOHA_ERR_LIMIT, "aloha_diff_backward: TF of %d entries exceeds LDS", a.tf_n);
  a.rgba = (float4 *)d_rgba;
  a.nsamples = (uint32_t *)d_nsamples;
  a.dldc = (const float4 *)d_dl_drgba;
  a.dtf = d_dtf;
  for (int k = 0; k < nb; ++k) a.dvox[k] = d_dvoxels ? d_dvoxels[k] : nullptr;
  DiffBins b;
  int max_dims[3];
  int64_t max_voxels = 0;
  const int64_t nbins = diff_layout(ctx, nb, brick_ids, a.dvox, b, max_dims, max_voxels);
  if (nbins > kMaxDiffBins)
    return fail(ctx, ALOHA_ERR_LIMIT, "aloha_diff_backward: %lld gradient bins", (long long)nbins);
  if (nbins > 0) {
    // slots of the waves (their total, read back, sizes the slot arrays);
    // bin counts and maxima start at zero
    const size_t words = 4 * (size_t)nbins + 1;
    if (words > ctx->diff_bins_cap) {
      if (ctx->diff_bins) HIP_CHECK(ctx, hipFree(ctx->diff_bins));
      ctx->diff_bins = nullptr;
      ctx->diff_bins_cap = 0;
      HIP_CHECK(ctx, dev_malloc((void **)&ctx->diff_bins, words * sizeof(uint32_t)));
      ctx->diff_bins_cap = words;
    }
    // first slot of each wave (an 8x8 tile of a view) and the slot counters
    // (8 bytes each), then each counter's list of waves with samples (the
    // waves of every kSlotCounters-th block of 4)
    const size_t nwaves = (size_t)V * ((H + 15) / 16) * ((W + 15) / 16) * 4;
    const size_t live_cap = (nwaves / 4 + kSlotCounters - 1) / kSlotCounters * 4;
    const size_t ray_bytes = 8 * (nwaves + kSlotCounters) + 4 * kSlotCounters * live_cap;
    if (ray_bytes > ctx->diff_rays_cap) {
      if (ctx->diff_rays) HIP_CHECK(ctx, hipFree(ctx->diff_rays));
      ctx->diff_rays = nullptr;
      ctx->diff_rays_cap = 0;
      HIP_CHECK(ctx, dev_malloc(&ctx->diff_rays, ray_bytes));
      ctx->diff_rays_cap = ray_bytes;
    }
    if (!ctx->h_total)
      HIP_CHECK(ctx, host_malloc((void **)&ctx->h_total, sizeof(uint64_t) * kSlotCounters));
    b.count = ctx->diff_bins;  // count | maxd | maxt | off (n + 1)
    b.maxd = b.count + nbins;
    b.maxt = b.maxd + nbins;
    b.off = b.maxt + nbins;
    b.nwaves = (int32_t)nwaves;
    b.wave_off = (unsigned long long *)ctx->diff_rays;
    b.slot_ctr = b.wave_off + nwaves;
    b.live = (int32_t *)(b.slot_ctr + kSlotCounters);
    b.live_cap = (uint32_t)live_cap;
    HIP_CHECK(ctx, hipMemsetAsync(b.count, 0, sizeof(uint32_t) * 3 * (size_t)nbins, ctx->stream));
    HIP_CHECK(ctx, hipMemsetAsync(b.slot_ctr, 0, sizeof(uint64_t) * kSlotCounters, ctx->stream));
    HIP_CHECK(ctx, launch_diff_rays(a, b, ctx->stream));
    HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_total, b.slot_ctr, sizeof(uint64_t) * kSlotCounters,
                                  hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    b.live_pre[0] = 0;
    for (int c = 0; c < kSlotCounters; ++c)
      b.live_pre[c + 1] = b.live_pre[c] + (uint32_t)(ctx->h_total[c] >> kLiveShift);
    // the counters' regions back to back
    size_t total = 0;
    for (int c = 0; c < kSlotCounters; ++c) {
      b.region[c] = total;
      total += (size_t)(ctx->h_total[c] & ((1ull << kLiveShift) - 1ull));
    }
    // (a counter's slots past 2^kLiveShift would spill into its wave count:
    // the counts then exceed the waves)
    if (total > ((size_t)1 << 31) || b.live_pre[kSlotCounters] > nwaves)
      return fail(ctx, ALOHA_ERR_LIMIT, "aloha_diff_backward: %zu sample slots", total);
    // per slot: resampled value (1 float4), run info (1), record (kRecWords),
    // run mark and list entry (one word each)
    constexpr size_t kSlotBytes = (2 + kRecWords) * sizeof(float4) + 2 * sizeof(uint32_t);
    if (total > ctx->diff_slots_cap) {
      if (ctx->diff_slots) HIP_CHECK(ctx, hipFree(ctx->diff_slots));
      ctx->diff_slots = nullptr;
      ctx->diff_slots_cap = 0;
      const size_t cap = total + total / 4 + 1024;
      HIP_CHECK(ctx, dev_malloc((void **)&ctx->diff_slots, cap * kSlotBytes));
      ctx->diff_slots_cap = cap;
    }
    const size_t tile_floats = max_voxels > 0 ? (size_t)nbins * kTileVox : 0;
    if (tile_floats > ctx->diff_tiles_cap) {
      if (ctx->diff_tiles) HIP_CHECK(ctx, hipFree(ctx->diff_tiles));
      ctx->diff_tiles = nullptr;
      ctx->diff_tiles_cap = 0;
      HIP_CHECK(ctx, dev_malloc((void **)&ctx->diff_tiles, tile_floats * sizeof(float)));
      ctx->diff_tiles_cap = tile_floats;
    }
    // per-bin TF partials, then the per-group sums of the reduction
    const size_t groups = ((size_t)nbins + kTfGroupBins - 1) / kTfGroupBins;
    const size_t part = d_dtf ? ((size_t)nbins + groups) * 4 * (size_t)a.tf_n : 0;
    if (part > ctx->diff_tf_part_cap) {
      if (ctx->diff_tf_part) HIP_CHECK(ctx, hipFree(ctx->diff_tf_part));
      ctx->diff_tf_part = nullptr;
      ctx->diff_tf_part_cap = 0;
      HIP_CHECK(ctx, dev_malloc((void **)&ctx->diff_tf_part, part * sizeof(double)));
      ctx->diff_tf_part_cap = part;
    }
    const size_t cap = ctx->diff_slots_cap;
    b.total = (uint32_t)total;
    b.cap = (uint32_t)cap;
    b.samples = ctx->diff_slots;
    b.run_info = b.samples + cap;
    b.records = b.run_info + cap;
    b.run_bin = (int32_t *)(b.records + kRecWords * cap);
    b.rec_list = (uint32_t *)(b.run_bin + cap);
    // no run marks but the heads the reverse walk writes
    HIP_CHECK(ctx, hipMemsetAsync(b.run_bin, 0xFF, sizeof(int32_t) * total, ctx->stream));
    b.tiles = ctx->diff_tiles;
    b.tf_part = ctx->diff_tf_part;
    HIP_CHECK(ctx, launch_diff_backward(a, b, ctx->stream));
    HIP_CHECK(ctx, launch_diff_bins(a, b, max_dims, ctx->stream));
  }
  return ALOHA_OK;
}


// ---------------------------------------------------------------------------
// batched form: many views of one rank table in one (graph-captured) launch
// ---------------------------------------------------------------------------
struct aloha_batch {
  aloha_ctx *ctx = nullptr;  // NULL once the context is destroyed
  uint64_t gen = 0;          // ctx->gen the launch arguments were made with
  BatchArgs b;
  float *d_cams = nullptr;
  int32_t *d_orders = nullptr, *d_rects = nullptr, *d_boxes = nullptr, *d_box_off = nullptr;
  uint32_t *d_counter = nullptr;
  int32_t *d_xoff = nullptr;
  int lanes = 2;
  int persistent_blocks = 0;
  int cus = 0;
  int P = 0;
  std::vector<int32_t> rects;  // V x P x 4 (the tile tables are made from them)
  hipStream_t stream = nullptr;  // capture / replay stream (not the null stream)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {
void batch_drop_graph(aloha_batch *bt) {
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  bt->exec = nullptr;
  bt->graph = nullptr;
}

bool lanes_valid(int n) { return n == 1 || n == 2 || n == 4 || n == 8; }

// Device resources of a batch (the context's device is current).
void batch_release(aloha_batch *bt) {
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  batch_drop_graph(bt);
  for (void *p : {(void *)bt->d_cams, (void *)bt->d_orders, (void *)bt->d_rects, (void *)bt->d_boxes,
                  (void *)bt->d_box_off, (void *)bt->d_counter, (void *)bt->d_xoff})
    if (p) (void)hipFree(p);
  bt->d_cams = nullptr, bt->d_orders = bt->d_rects = bt->d_boxes = bt->d_box_off = nullptr;
  bt->d_counter = nullptr;
  bt->d_xoff = nullptr;
  if (bt->ev_in) (void)hipEventDestroy(bt->ev_in);
  if (bt->ev_out) (void)hipEventDestroy(bt->ev_out);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  bt->ev_in = bt->ev_out = nullptr;
  bt->stream = nullptr;
}

// The context's TF or brick table moved (or the TF changed size) since the
// batch's launch arguments were made: take the current ones and drop a
// captured graph, which holds the old ones (capture again if there was one).
int batch_refresh(aloha_batch *bt);

// The work items of tiles (1 << wl) x (64 >> wl): per view the tile
// bounding box of its rectangles and the prefix of the active items, and the
// persistent schedule's per-XCD parts (part x of view v is its items
// [n x / kXcds, n (x + 1) / kXcds), n = the view's item count); copied to
// the batch's device tables once the batch stream is idle.
int batch_tables(aloha_batch *bt, int wl) {
  aloha_ctx *ctx = bt->ctx;
  const int V = bt->b.V, P = bt->P, W = bt->b.m.W, H = bt->b.m.H;
  const int tw = 1 << wl, th = 64 >> wl;
  std::vector<int32_t> boxes((size_t)V * 4, 0), box_off((size_t)V + 1, 0);
  for (int v = 0; v < V; ++v) {
    int x0 = W, y0 = H, x1 = 0, y1 = 0;
    for (int s = 0; s < P; ++s) {
      const int32_t *rc = &bt->rects[((size_t)v * P + s) * 4];
      if (rc[0] >= rc[2] || rc[1] >= rc[3]) continue;
      x0 = std::min(x0, (int)rc[0]), y0 = std::min(y0, (int)rc[1]);
      x1 = std::max(x1, (int)rc[2]), y1 = std::max(y1, (int)rc[3]);
    }
    int32_t *bx = &boxes[(size_t)v * 4];
    if (x0 < x1 && y0 < y1) {
      bx[0] = x0 / tw, bx[1] = y0 / th;
      bx[2] = (x1 + tw - 1) / tw - bx[0];
      bx[3] = (y1 + th - 1) / th - bx[1];
    }
    const int64_t next = (int64_t)box_off[v] + (int64_t)bx[2] * bx[3];
    if (next > INT32_MAX) return fail(ctx, ALOHA_ERR_LIMIT, "batch: too many tiles");
    box_off[v + 1] = (int32_t)next;
  }
  std::vector<int32_t> xoff((size_t)kXcds * (V + 1), 0);
  for (int x = 0; x < kXcds; ++x)
    for (int v = 0; v < V; ++v) {
      const int64_t n = (int64_t)box_off[v + 1] - box_off[v];
      const int64_t part = n * (x + 1) / kXcds - n * x / kXcds;
      xoff[(size_t)x * (V + 1) + v + 1] = xoff[(size_t)x * (V + 1) + v] + (int32_t)part;
    }
  if (bt->stream) HIP_CHECK(ctx, hipStreamSynchronize(bt->stream));
  HIP_CHECK(ctx, hipMemcpy(bt->d_boxes, boxes.data(), sizeof(int32_t) * boxes.size(),
                           hipMemcpyHostToDevice));
  HIP_CHECK(ctx, hipMemcpy(bt->d_box_off, box_off.data(), sizeof(int32_t) * box_off.size(),
                           hipMemcpyHostToDevice));
  HIP_CHECK(ctx, hipMemcpy(bt->d_xoff, xoff.data(), sizeof(int32_t) * xoff.size(),
                           hipMemcpyHostToDevice));
  bt->b.total_active = box_off[V];
  bt->b.tile_wl = wl;
  return ALOHA_OK;
}
}  // namespace

namespace aloha {
void batch_detach(aloha_batch *bt) {
  batch_release(bt);
  bt->ctx = nullptr;
}
}  // namespace aloha

extern "C" int aloha_batch_create(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                  int32_t H, int32_t V, const float *cams, int32_t P,
                                  const int32_t *rank_offsets, const int32_t *brick_ids,
                                  const int32_t *orders, float *d_out_rgba,
                                  uint32_t *d_out_rgba8, uint32_t *d_nsamples,
                                  aloha_batch **out) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_create: ctx is NULL");
  if (!out || !cams || !orders || V < 1)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_batch_create: NULL argument or V < 1");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // camera of view 0 stands in for build_march's camera check; views are validated below
  const float saved[12] = {ctx->cam[0], ctx->cam[1], ctx->cam[2], ctx->cam[3], ctx->cam[4], ctx->cam[5],
                           ctx->cam[6], ctx->cam[7], ctx->cam[8], ctx->cam[9], ctx->cam[10], ctx->cam[11]};
  const bool had = ctx->has_cam;
  for (int i = 0; i < 12 * V; ++i)
    if (!std::isfinite(cams[i])) return fail(ctx, ALOHA_ERR_ARG, "aloha_batch_create: non-finite camera");
  std::memcpy(ctx->cam, cams, sizeof saved);
  ctx->has_cam = true;
  MarchArgs m;
  int rc = build_march(ctx, params, W, H, P, rank_offsets, brick_ids, m);
  std::memcpy(ctx->cam, saved, sizeof saved);
  ctx->has_cam = had;
  if (rc) return rc;
  if (params->accum_id != 0 || (params->flags & (ALOHA_FLAG_HEATMAP | ALOHA_FLAG_CROSSHAIR)))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_batch_create: accum_id 0 and no heat map / crosshair only");
  if ((int64_t)V * W * H > (int64_t)1 << 31)
    return fail(ctx, ALOHA_ERR_LIMIT, "aloha_batch_create: %d views of %dx%d too large", V, W, H);
  std::vector<int32_t> rects((size_t)V * P * 4);
  for (int v = 0; v < V; ++v) {
    std::vector<char> seen(P, 0);
    for (int k = 0; k < P; ++k) {
      const int o = orders[(size_t)v * P + k];
      if (o < 0 || o >= P || seen[o])
        return fail(ctx, ALOHA_ERR_ARG, "aloha_batch_create: order of view %d is not a permutation", v);
      seen[o] = 1;
    }
    if (int rc2 = check_steps(ctx, cams + 12 * v, ctx->bricks, brick_ids, rank_offsets[P], params->dt))
      return rc2;
    for (int s = 0; s < P; ++s)
      rank_rect(cams + 12 * v, ctx->bricks, brick_ids + rank_offsets[s],
                rank_offsets[s + 1] - rank_offsets[s], W, H, &rects[((size_t)v * P + s) * 4]);
  }
  if ((int64_t)V * H > INT32_MAX) return fail(ctx, ALOHA_ERR_LIMIT, "aloha_batch_create: too many rows");
  auto *bt = new aloha_batch;
  bt->ctx = ctx;
  std::memset(&bt->b, 0, sizeof bt->b);
  bt->b.m = m;
  bt->b.m.rgba = (float4 *)d_out_rgba;
  bt->b.m.rgba8 = d_out_rgba8;
  bt->b.m.nsamples = d_nsamples;
  bt->b.V = V;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, ctx->device);
  bt->cus = e == hipSuccess ? prop.multiProcessorCount : 256;
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_cams, sizeof(float) * 12 * (size_t)V);
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_orders, sizeof(int32_t) * (size_t)V * P);
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_rects, sizeof(int32_t) * rects.size());
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_boxes, sizeof(int32_t) * 4 * (size_t)V);
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_box_off, sizeof(int32_t) * ((size_t)V + 1));
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_counter, sizeof(uint32_t) * kXcds);
  if (e == hipSuccess) e = dev_malloc((void **)&bt->d_xoff, sizeof(int32_t) * kXcds * ((size_t)V + 1));
  if (e == hipSuccess) e = hipMemcpy(bt->d_cams, cams, sizeof(float) * 12 * (size_t)V, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(bt->d_orders, orders, sizeof(int32_t) * (size_t)V * P, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(bt->d_rects, rects.data(), sizeof(int32_t) * rects.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    aloha_batch_destroy(bt);
    return fail(ctx, e == hipErrorOutOfMemory ? ALOHA_ERR_ALLOC : ALOHA_ERR_HIP,
                "aloha_batch_create: %s", hipGetErrorString(e));
  }
  bt->b.cams = bt->d_cams;
  bt->b.orders = bt->d_orders;
  bt->b.rects = bt->d_rects;
  bt->b.boxes = bt->d_boxes;
  bt->b.box_off = bt->d_box_off;
  bt->b.counter = bt->d_counter;
  bt->b.xoff = bt->d_xoff;
  bt->b.fill_blocks = (int)std::min<int64_t>((int64_t)V * H, 1024);
  bt->P = P;
  bt->rects = std::move(rects);
  if (int rc = batch_tables(bt, 4)) {  // 16x4 tiles (the single fused frame's)
    aloha_batch_destroy(bt);
    return rc;
  }
  bt->lanes = ctx->fused_lanes;
  bt->gen = ctx->gen;
  ctx->batches.push_back(bt);
  *out = bt;
  return ALOHA_OK;
}

extern "C" int aloha_batch_destroy(aloha_batch *bt) {
  if (!bt) return ALOHA_OK;
  if (aloha_ctx *ctx = bt->ctx) {
    (void)hipSetDevice(ctx->device);
    batch_release(bt);
    auto &v = ctx->batches;
    v.erase(std::remove(v.begin(), v.end(), bt), v.end());
  }
  delete bt;
  return ALOHA_OK;
}

extern "C" int aloha_batch_set_schedule(aloha_batch *bt, int32_t lanes_per_ray,
                                        int32_t persistent_blocks) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_set_schedule: NULL");
  if (!bt->ctx) return fail(nullptr, ALOHA_ERR_STATE, "aloha_batch_set_schedule: context destroyed");
  if (!lanes_valid(lanes_per_ray) || persistent_blocks < 0 || persistent_blocks > (1 << 20))
    return fail(bt->ctx, ALOHA_ERR_ARG,
                "aloha_batch_set_schedule: lanes per ray 1, 2, 4 or 8; blocks >= 0");
  if (lanes_per_ray != bt->lanes || persistent_blocks != bt->persistent_blocks) batch_drop_graph(bt);
  bt->lanes = lanes_per_ray;
  bt->persistent_blocks = persistent_blocks;
  return ALOHA_OK;
}

extern "C" int aloha_batch_get_schedule(const aloha_batch *bt, int32_t *lanes_per_ray,
                                        int32_t *persistent_blocks, int32_t *captured) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_get_schedule: NULL");
  if (lanes_per_ray) *lanes_per_ray = bt->lanes;
  if (persistent_blocks) *persistent_blocks = bt->persistent_blocks;
  if (captured) *captured = bt->exec != nullptr;
  return ALOHA_OK;
}

extern "C" int aloha_batch_set_tile(aloha_batch *bt, int32_t tile_w) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_set_tile: NULL");
  if (!bt->ctx) return fail(nullptr, ALOHA_ERR_STATE, "aloha_batch_set_tile: context destroyed");
  if (tile_w != 8 && tile_w != 16 && tile_w != 32)
    return fail(bt->ctx, ALOHA_ERR_ARG, "aloha_batch_set_tile: tile width 8, 16 or 32");
  const int wl = tile_w == 8 ? 3 : tile_w == 16 ? 4 : 5;
  if (wl == bt->b.tile_wl) return ALOHA_OK;
  HIP_CHECK(bt->ctx, hipSetDevice(bt->ctx->device));
  if (bt->stream) HIP_CHECK(bt->ctx, hipStreamSynchronize(bt->stream));
  batch_drop_graph(bt);
  return batch_tables(bt, wl);
}

extern "C" int aloha_batch_get_tile(const aloha_batch *bt, int32_t *tile_w) {
  if (!bt || !tile_w) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_get_tile: NULL");
  *tile_w = 1 << bt->b.tile_wl;
  return ALOHA_OK;
}

// Launch on the batch stream, ordered after earlier work on the ctx stream
// and before later work on it.
extern "C" int aloha_batch_launch(aloha_batch *bt) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_launch: NULL");
  if (!bt->ctx) return fail(nullptr, ALOHA_ERR_STATE, "aloha_batch_launch: context destroyed");
  aloha_ctx *ctx = bt->ctx;
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (bt->gen != ctx->gen)
    if (int rc = batch_refresh(bt)) return rc;
  HIP_CHECK(ctx, hipEventRecord(bt->ev_in, ctx->stream));
  HIP_CHECK(ctx, hipStreamWaitEvent(bt->stream, bt->ev_in, 0));
  if (bt->exec)
    HIP_CHECK(ctx, hipGraphLaunch(bt->exec, bt->stream));
  else
    HIP_CHECK(ctx, launch_batch(bt->b, bt->lanes, bt->persistent_blocks, bt->stream));
  HIP_CHECK(ctx, hipEventRecord(bt->ev_out, bt->stream));
  HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, bt->ev_out, 0));
  return ALOHA_OK;
}

extern "C" int aloha_batch_capture(aloha_batch *bt) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_capture: NULL");
  if (!bt->ctx) return fail(nullptr, ALOHA_ERR_STATE, "aloha_batch_capture: context destroyed");
  aloha_ctx *ctx = bt->ctx;
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (bt->gen != ctx->gen)
    if (int rc = batch_refresh(bt)) return rc;
  batch_drop_graph(bt);
  HIP_CHECK(ctx, hipStreamSynchronize(bt->stream));
  HIP_CHECK(ctx, hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch(bt->b, bt->lanes, bt->persistent_blocks, bt->stream);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(bt->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(ctx, ALOHA_ERR_HIP, "aloha_batch_capture: %s",
                hipGetErrorString(e != hipSuccess ? e : e2));
  }
  bt->graph = g;
  e = hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    batch_drop_graph(bt);
    return fail(ctx, ALOHA_ERR_HIP, "aloha_batch_capture: instantiate: %s", hipGetErrorString(e));
  }
  return ALOHA_OK;
}

namespace {
int batch_refresh(aloha_batch *bt) {
  aloha_ctx *ctx = bt->ctx;
  if (!ctx->d_tf || ctx->tf_n < 1) return fail(ctx, ALOHA_ERR_STATE, "batch: no transfer function set");
  if (int rc = sync_bricks(ctx)) return rc;
  const bool was_captured = bt->exec != nullptr;
  // the old graph may still run on the batch stream
  HIP_CHECK(ctx, hipStreamSynchronize(bt->stream));
  batch_drop_graph(bt);
  bt->b.m.tf = ctx->d_tf;
  bt->b.m.tf_n = ctx->tf_n;
  bt->b.m.bricks = ctx->d_bricks;
  bt->gen = ctx->gen;
  return was_captured ? aloha_batch_capture(bt) : ALOHA_OK;
}
}  // namespace

// Times every schedule (tiles 8x8, 16x4, 32x2 x 1/2/4/8 lanes per ray x {one
// block per item, persistent with 2/4/8 blocks per CU}) over `reps` launches
// each and keeps the fastest.  The schedule never changes the samples or
// their arithmetic, so the images do not depend on the choice.  Synchronous.
extern "C" int aloha_batch_autotune(aloha_batch *bt, int32_t reps, float *best_ms,
                                    float *table, int32_t table_cap, int32_t *table_n) {
  if (!bt) return fail(nullptr, ALOHA_ERR_ARG, "aloha_batch_autotune: NULL");
  if (!bt->ctx) return fail(nullptr, ALOHA_ERR_STATE, "aloha_batch_autotune: context destroyed");
  aloha_ctx *ctx = bt->ctx;
  if (reps < 1) return fail(ctx, ALOHA_ERR_ARG, "aloha_batch_autotune: reps < 1");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (bt->gen != ctx->gen)
    if (int rc = batch_refresh(bt)) return rc;
  batch_drop_graph(bt);
  hipEvent_t e0, e1;
  HIP_CHECK(ctx, hipEventCreate(&e0));
  HIP_CHECK(ctx, hipEventCreate(&e1));
  float best = 1e30f;
  int best_lanes = bt->lanes, best_blocks = 0, best_wl = bt->b.tile_wl, n = 0;
  const int blocks_opts[4] = {0, 2, 4, 8};
  const int lane_opts[4] = {1, 2, 4, 8};
  for (int wl = 3; wl <= 5; ++wl) {
    if (int rc = batch_tables(bt, wl)) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return rc;
    }
    for (int lanes : lane_opts)
    for (int bo : blocks_opts) {
      bt->lanes = lanes;
      bt->persistent_blocks = bo ? bo * bt->cus : 0;
      hipError_t e = launch_batch(bt->b, bt->lanes, bt->persistent_blocks, bt->stream);  // warm
      if (e == hipSuccess) e = hipEventRecord(e0, bt->stream);
      for (int r = 0; e == hipSuccess && r < reps; ++r)
        e = launch_batch(bt->b, bt->lanes, bt->persistent_blocks, bt->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, bt->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return fail(ctx, ALOHA_ERR_HIP, "aloha_batch_autotune: %s", hipGetErrorString(e));
      }
      ms /= (float)reps;
      if (table && n < table_cap) {
        table[4 * n] = (float)(1 << wl);
        table[4 * n + 1] = (float)lanes;
        table[4 * n + 2] = (float)bt->persistent_blocks;
        table[4 * n + 3] = ms;
      }
      ++n;
      if (ms < best) {
        best = ms;
        best_lanes = lanes;
        best_blocks = bt->persistent_blocks;
        best_wl = wl;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  bt->lanes = best_lanes;
  bt->persistent_blocks = best_blocks;
  if (int rc = batch_tables(bt, best_wl)) return rc;
  if (best_ms) *best_ms = best;
  if (table_n) *table_n = n;
  return ALOHA_OK;
}
