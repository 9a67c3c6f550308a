// host_mesh.h -- mesh handle and BVH layout shared by host_mesh.cpp and the
// device context.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "aloha_amd.h"

struct MeshCluster {
  int64_t first, last;  // triangle range [first, last)
  aloha_box3f bounds;   // split domain (clipped at the split planes)
};

struct aloha_mesh {
  std::vector<float> verts;    // 3 floats per vertex
  std::vector<int32_t> tris;   // 3 vertex ids per triangle
  std::vector<MeshCluster> clusters;
  aloha_box3f bounds;          // bounds of all vertices
};

namespace aloha {

// The BVH is built binary (binned SAH) and collapsed to four children per
// node, the layout the traversal kernel reads: 128 bytes, so one node fetch
// (eight 16-byte words) tests four boxes.
//   words 0-5: lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4]
//   word 6:    ref[4];  word 7: zero
// A child ref >= 0 is a node index; a leaf is ~(first << kLeafBits |
// (count - 1)) over the BVH-ordered triangle array; an unused child slot is
// kEmptyChild (its box empty: lo = +FLT_MAX, hi = -FLT_MAX).
constexpr int kLeafBits = 3;
constexpr int kMaxLeaf = 1 << kLeafBits;  // triangles per leaf, at most
constexpr int kMaxBvhDepth = 32;          // leaf depth bound of the binary build
constexpr int32_t kEmptyChild = 0x7fffffff;
struct TriNode4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int32_t ref[4];
  int32_t zero[4];
};
static_assert(sizeof(TriNode4) == 128, "BVH node layout");

struct TriBvh {
  std::vector<TriNode4> nodes;  // nodes, depth-first (a node's children after it)
  std::vector<uint32_t> order;  // BVH triangle slot -> input triangle
  int32_t root = 0;             // node 0, or a leaf ref when there is one leaf
  int32_t depth = 0;            // deepest leaf of the 4-wide tree (root = 0)
  int32_t depth2 = 0;           // deepest leaf of the binary build
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // the triangles' bounds
  float pad = 0.f;                              // the node boxes' widening
};

// Binned-SAH BVH over ntris triangles given as 9 floats each (three
// vertices).  Node boxes are widened by `pad` on every side so that float
// slab tests never cull a box whose triangle the intersection test hits.
// Returns false (with a message) when the mesh is too large to encode.
bool build_tri_bvh(const float *tri9, int64_t ntris, float pad, TriBvh &out, std::string &err);

// One rank's geometry on its way to HBM: the listed clusters' triangles
// flattened to three private vertices each (list order = primitive id),
// their BVH, and the triangle array in leaf order (12 floats per triangle:
// three vertices whose w hold the primitive id and colour id bits).
struct RankBuild {
  std::vector<float> tri9;
  std::vector<uint32_t> color;
  TriBvh bvh;
  std::vector<float> dev_tris;
  std::string err;
  bool ok = true;
};
void flatten(const aloha_mesh *m, const int32_t *clusters, const uint32_t *colors, int n,
             RankBuild &rb);
void build_rank(RankBuild &rb);

}  // namespace aloha
