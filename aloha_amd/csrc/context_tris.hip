// context_tris.hip -- the surface path's entry points on a device context:
// triangle upload (flatten + BVH build + HBM copy) and the render launches.
// Replaces Renderer's TriMesh set-up (waikikiRenderer.cpp:390-593: .tri
// load, flatten to three private vertices per triangle, one OptiX BLAS per
// cluster under a TLAS) and the mesh frame (:1028, :1070-1079).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "context.h"
#include "host_mesh.h"

using namespace aloha;

namespace {

constexpr int32_t kEmptyRoot = INT32_MIN;  // the kernels' marker of a rank with no triangles
// a walk pushes at most 3 children per 4-wide level, and a 4-wide leaf is no
// deeper than its binary one
static_assert(kTriStack == 3 * kMaxBvhDepth, "traversal stack must cover the deepest leaf");

int geom_of(const aloha_ctx *ctx, int32_t id) {
  return id >= 0 && id < (int32_t)ctx->geoms.size() ? id : -1;
}

}  // namespace

extern "C" int aloha_tris_upload(aloha_ctx *ctx, const aloha_mesh *mesh, int32_t P,
                                 const int32_t *rank_offsets, const int32_t *cluster_ids,
                                 const uint32_t *color_ids, int32_t *geom_ids) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_tris_upload: ctx is NULL");
  if (!mesh || !rank_offsets || !geom_ids)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_tris_upload: NULL argument");
  if (P < 1 || P > ALOHA_MAX_RANKS)
    return fail(ctx, ALOHA_ERR_LIMIT, "aloha_tris_upload: P=%d outside [1,%d]", P, ALOHA_MAX_RANKS);
  if (rank_offsets[0] != 0) return fail(ctx, ALOHA_ERR_ARG, "aloha_tris_upload: rank_offsets[0] != 0");
  for (int r = 0; r < P; ++r)
    if (rank_offsets[r + 1] < rank_offsets[r])
      return fail(ctx, ALOHA_ERR_ARG, "aloha_tris_upload: rank_offsets not monotone at %d", r);
  const int nl = rank_offsets[P];
  if (nl > 0 && (!cluster_ids || !color_ids))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_tris_upload: cluster_ids / color_ids NULL");
  // an unsplit mesh is one cluster holding every triangle
  aloha_mesh whole;
  const aloha_mesh *m = mesh;
  if (mesh->clusters.empty()) {
    whole.verts = mesh->verts;
    whole.tris = mesh->tris;
    whole.bounds = mesh->bounds;
    whole.clusters.push_back(MeshCluster{0, (int64_t)mesh->tris.size() / 3, mesh->bounds});
    m = &whole;
  }
  for (int k = 0; k < nl; ++k)
    if (cluster_ids[k] < 0 || cluster_ids[k] >= (int)m->clusters.size())
      return fail(ctx, ALOHA_ERR_ARG, "aloha_tris_upload: cluster id %d of %d", cluster_ids[k],
                  (int)m->clusters.size());
  HIP_CHECK(ctx, hipSetDevice(ctx->device));

  std::vector<RankBuild> rb(P);
  for (int r = 0; r < P; ++r)
    flatten(m, cluster_ids + rank_offsets[r], color_ids + rank_offsets[r],
            rank_offsets[r + 1] - rank_offsets[r], rb[r]);
  {
    const int nth = (int)std::max(1u, std::min<unsigned>(P, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t)
      th.emplace_back([&, t] {
        for (int r = t; r < P; r += nth) build_rank(rb[r]);
      });
    for (auto &x : th) x.join();
  }
  for (int r = 0; r < P; ++r)
    if (!rb[r].ok) return fail(ctx, ALOHA_ERR_LIMIT, "aloha_tris_upload: rank %d: %s", r, rb[r].err.c_str());

  const size_t first_id = ctx->geoms.size();
  for (int r = 0; r < P; ++r) {
    TriGeom g;
    g.ntris = (int64_t)rb[r].color.size();
    if (g.ntris == 0) {
      g.root = kEmptyRoot;
    } else {
      const TriBvh &b = rb[r].bvh;
      if (3 * b.depth > kTriStack)
        return fail(ctx, ALOHA_ERR_LIMIT, "aloha_tris_upload: BVH depth %d exceeds %d", b.depth,
                    kTriStack / 3);
      g.root = b.root;
      g.depth = b.depth;
      for (int k = 0; k < 3; ++k) {
        g.box.lower[k] = b.lo[k] - b.pad;
        g.box.upper[k] = b.hi[k] + b.pad;
      }
      g.nnodes = (int64_t)b.nodes.size();
      const size_t nb = std::max<size_t>(1, b.nodes.size()) * sizeof(TriNode4);
      const size_t tb = rb[r].dev_tris.size() * sizeof(float);
      hipError_t e = dev_malloc((void **)&g.d_nodes, nb);
      if (e == hipSuccess) e = dev_malloc((void **)&g.d_tris, tb);
      if (e == hipSuccess && !b.nodes.empty())
        e = hipMemcpy(g.d_nodes, b.nodes.data(), b.nodes.size() * sizeof(TriNode4),
                      hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(g.d_tris, rb[r].dev_tris.data(), tb, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        if (g.d_nodes) hipFree(g.d_nodes);
        if (g.d_tris) hipFree(g.d_tris);
        // the ranks uploaded by this call go too: a failed call adds nothing
        for (size_t k = first_id; k < ctx->geoms.size(); ++k) {
          if (ctx->geoms[k].d_nodes) hipFree(ctx->geoms[k].d_nodes);
          if (ctx->geoms[k].d_tris) hipFree(ctx->geoms[k].d_tris);
        }
        ctx->geoms.resize(first_id);
        return fail(ctx, e == hipErrorOutOfMemory ? ALOHA_ERR_ALLOC : ALOHA_ERR_HIP,
                    "aloha_tris_upload: %s", hipGetErrorString(e));
      }
      g.bytes = nb + tb;
    }
    ctx->geoms.push_back(g);
    geom_ids[r] = (int32_t)(first_id + r);
  }
  return ALOHA_OK;
}

extern "C" int aloha_tris_info(const aloha_ctx *ctx, int32_t geom, int64_t *ntris, int64_t *nnodes,
                               int32_t *depth, uint64_t *bytes) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_tris_info: ctx is NULL");
  const int g = geom_of(ctx, geom);
  if (g < 0) return fail(const_cast<aloha_ctx *>(ctx), ALOHA_ERR_ARG, "aloha_tris_info: no geometry %d", geom);
  const TriGeom &t = ctx->geoms[g];
  if (ntris) *ntris = t.ntris;
  if (nnodes) *nnodes = t.nnodes;
  if (depth) *depth = t.depth;
  if (bytes) *bytes = t.bytes;
  return ALOHA_OK;
}

extern "C" int aloha_tris_export(const aloha_ctx *ctx, int32_t geom, void *nodes, float *tris,
                                 int32_t *root) {
  aloha_ctx *c = const_cast<aloha_ctx *>(ctx);
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_tris_export: ctx is NULL");
  const int g = geom_of(ctx, geom);
  if (g < 0) return fail(c, ALOHA_ERR_ARG, "aloha_tris_export: no geometry %d", geom);
  const TriGeom &t = ctx->geoms[g];
  HIP_CHECK(c, hipSetDevice(ctx->device));
  HIP_CHECK(c, hipStreamSynchronize(ctx->stream));
  if (root) *root = t.root;
  if (nodes && t.nnodes)
    HIP_CHECK(c, hipMemcpy(nodes, t.d_nodes, t.nnodes * sizeof(TriNode4), hipMemcpyDeviceToHost));
  if (tris && t.ntris)
    HIP_CHECK(c, hipMemcpy(tris, t.d_tris, t.ntris * 48, hipMemcpyDeviceToHost));
  return ALOHA_OK;
}

namespace {

int build_tris(aloha_ctx *ctx, const aloha_render_params *p, int32_t W, int32_t H, int32_t P,
               const int32_t *geom_ids, TrisArgs &a) {
  if (int rc = check_params(ctx, p)) return rc;
  if (W < 1 || H < 1 || (int64_t)W * H > (int64_t)1 << 30)
    return fail(ctx, ALOHA_ERR_ARG, "render_tris: image %dx%d out of range", W, H);
  if (P < 1 || P > ALOHA_MAX_RANKS)
    return fail(ctx, ALOHA_ERR_LIMIT, "render_tris: P=%d outside [1,%d]", P, ALOHA_MAX_RANKS);
  if (!geom_ids) return fail(ctx, ALOHA_ERR_ARG, "render_tris: geom_ids is NULL");
  if (!ctx->has_cam) return fail(ctx, ALOHA_ERR_STATE, "render_tris: no camera set");
  std::memset(&a, 0, sizeof a);
  for (int r = 0; r < P; ++r) {
    const int g = geom_of(ctx, geom_ids[r]);
    if (g < 0) return fail(ctx, ALOHA_ERR_ARG, "render_tris: geometry id %d not uploaded", geom_ids[r]);
    const TriGeom &t = ctx->geoms[g];
    a.geo[r].nodes = t.d_nodes;
    a.geo[r].tris = t.d_tris;
    a.geo[r].root = t.root;
    // the pixels whose ray can reach the geometry's (widened) bounds: every
    // box of its BVH lies inside them, so the walk of any other pixel finds
    // nothing
    if (t.root == kEmptyRoot) {
      for (int k = 0; k < 4; ++k) a.rect[r][k] = 0;
    } else {
      project_boxes(ctx->cam, &t.box, 1, W, H, a.rect[r]);
    }
  }
  std::memcpy(a.org, ctx->cam, sizeof a.org);
  std::memcpy(a.dir00, ctx->cam + 3, sizeof a.dir00);
  std::memcpy(a.du, ctx->cam + 6, sizeof a.du);
  std::memcpy(a.dv, ctx->cam + 9, sizeof a.dv);
  a.max_depth = p->max_depth;
  a.heat_scale = p->heat_map_scale;
  a.heat_clock = p->heat_clock;
  a.W = W;
  a.H = H;
  a.spp = p->spp;
  a.accum_id = p->accum_id;
  a.island_id = p->island_id;
  a.num_islands = p->num_islands;
  a.flags = p->flags;
  a.P = P;
  return ALOHA_OK;
}

}  // namespace

extern "C" int aloha_render_tris(aloha_ctx *ctx, const aloha_render_params *params, int32_t W,
                                 int32_t H, int32_t P, const int32_t *geom_ids,
                                 const aloha_images *images, uint32_t *d_stats) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_tris: ctx is NULL");
  if (!images) return fail(ctx, ALOHA_ERR_ARG, "aloha_render_tris: images is NULL");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  TrisArgs a;
  if (int rc = build_tris(ctx, params, W, H, P, geom_ids, a)) return rc;
  if (params->accum_id > 0 && !images->accum)
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_tris: accum_id > 0 needs the accum buffer");
  a.rgba = (float4 *)images->rgba;
  a.rgba8 = images->rgba8;
  a.depth = images->depth;
  a.accum = (float4 *)images->accum;
  a.stats = d_stats;
  HIP_CHECK(ctx, launch_tris(a, false, ctx->stream));
  return ALOHA_OK;
}

extern "C" int aloha_render_tris_composite(aloha_ctx *ctx, const aloha_render_params *params,
                                           int32_t W, int32_t H, int32_t P,
                                           const int32_t *geom_ids, uint32_t *d_out_rgba8,
                                           float *d_out_depth, uint32_t *d_stats) {
  if (!ctx) return fail(nullptr, ALOHA_ERR_ARG, "aloha_render_tris_composite: ctx is NULL");
  HIP_CHECK(ctx, hipSetDevice(ctx->device));
  TrisArgs a;
  if (int rc = build_tris(ctx, params, W, H, P, geom_ids, a)) return rc;
  if (params->accum_id != 0 || (params->flags & ALOHA_FLAG_HEATMAP))
    return fail(ctx, ALOHA_ERR_ARG, "aloha_render_tris_composite: accum_id 0 and no heat map only");
  a.rgba8 = d_out_rgba8;
  a.depth = d_out_depth;
  a.stats = d_stats;
  HIP_CHECK(ctx, launch_tris(a, true, ctx->stream));
  return ALOHA_OK;
}
