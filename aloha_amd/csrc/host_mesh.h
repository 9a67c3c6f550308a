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

// Binary BVH node as the traversal kernel reads it: 64 bytes, four 16-byte
// words holding both children's boxes, so one node fetch tests two boxes.
//   word 0: left  {lo.x, hi.x, lo.y, hi.y}
//   word 1: right {lo.x, hi.x, lo.y, hi.y}
//   word 2: {left lo.z, left hi.z, right lo.z, right hi.z}
//   word 3: {left ref, right ref, 0, 0}
// A child ref >= 0 is an internal node index; a leaf is
// ~(first << kLeafBits | (count - 1)) over the BVH-ordered triangle array.
constexpr int kLeafBits = 3;
constexpr int kMaxLeaf = 1 << kLeafBits;  // triangles per leaf, at most
constexpr int kMaxBvhDepth = 32;          // leaf depth bound = traversal stack size
struct TriNode {
  float l_xy[4], r_xy[4], lr_z[4];
  int32_t ref[4];
};
static_assert(sizeof(TriNode) == 64, "BVH node layout");

struct TriBvh {
  std::vector<TriNode> nodes;   // internal nodes, depth-first (left child next)
  std::vector<uint32_t> order;  // BVH triangle slot -> input triangle
  int32_t root = 0;             // node 0, or a leaf ref when there is one leaf
  int32_t depth = 0;            // deepest leaf (root = 0)
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
