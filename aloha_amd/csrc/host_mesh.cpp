// host_mesh.cpp -- triangle meshes for the surface path, on the host:
//
//   aloha_mesh_synth_spheres   seeded displaced UV spheres (benchmark input)
//   aloha_mesh_split           MeshSplitter, median strategy (chopSuey.cpp:48-229)
//   aloha_tri_write / _read    the .tri format (chopSuey.cpp:200-218,
//                              waikikiRenderer.cpp:394-451)
//   build_tri_bvh              binned-SAH BVH over one rank's flattened
//                              triangles (replaces OptiX's accel build,
//                              waikikiRenderer.cpp:540-593)
//
// Two orderings the reference leaves to std::sort / std::partition (both
// unstable) are made deterministic here: primitive references sort by
// (centroid, primitive id) and triangles are partitioned stably.
#include "host_mesh.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <thread>

#include "internal.h"

using namespace aloha;

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

void box_empty(aloha_box3f &b) {
  for (int a = 0; a < 3; ++a) {
    b.lower[a] = 1e30f;
    b.upper[a] = -1e30f;
  }
}
void box_extend(aloha_box3f &b, const float *p) {
  for (int a = 0; a < 3; ++a) {
    b.lower[a] = std::min(b.lower[a], p[a]);
    b.upper[a] = std::max(b.upper[a], p[a]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// mesh handle
// ---------------------------------------------------------------------------
extern "C" int aloha_mesh_create(const float *verts, int64_t nverts, const int32_t *tris,
                                 int64_t ntris, aloha_mesh **out) {
  if (!out || nverts < 0 || ntris < 0 || (nverts && !verts) || (ntris && !tris))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_create: bad argument");
  for (int64_t i = 0; i < 3 * ntris; ++i)
    if (tris[i] < 0 || tris[i] >= nverts)
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_create: index %d out of range", tris[i]);
  auto *m = new aloha_mesh;
  m->verts.assign(verts, verts + 3 * nverts);
  m->tris.assign(tris, tris + 3 * ntris);
  box_empty(m->bounds);
  for (int64_t i = 0; i < nverts; ++i) box_extend(m->bounds, &m->verts[3 * i]);
  *out = m;
  return ALOHA_OK;
}

extern "C" void aloha_mesh_destroy(aloha_mesh *m) { delete m; }

extern "C" int aloha_mesh_info(const aloha_mesh *m, int64_t *nverts, int64_t *ntris,
                               aloha_box3f *bounds, int32_t *nclusters) {
  if (!m) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_info: NULL");
  if (nverts) *nverts = (int64_t)m->verts.size() / 3;
  if (ntris) *ntris = (int64_t)m->tris.size() / 3;
  if (bounds) *bounds = m->bounds;
  if (nclusters) *nclusters = (int32_t)m->clusters.size();
  return ALOHA_OK;
}

extern "C" int aloha_mesh_get(const aloha_mesh *m, float *verts, int32_t *tris) {
  if (!m) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_get: NULL");
  if (verts) std::memcpy(verts, m->verts.data(), m->verts.size() * sizeof(float));
  if (tris) std::memcpy(tris, m->tris.data(), m->tris.size() * sizeof(int32_t));
  return ALOHA_OK;
}

extern "C" int aloha_mesh_clusters(const aloha_mesh *m, int64_t *ranges, aloha_box3f *bounds,
                                   int32_t cap, int32_t *count) {
  if (!m || !count) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_clusters: NULL");
  *count = (int32_t)m->clusters.size();
  if ((int)m->clusters.size() > cap)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_mesh_clusters: cap too small");
  for (size_t i = 0; i < m->clusters.size(); ++i) {
    if (ranges) {
      ranges[2 * i] = m->clusters[i].first;
      ranges[2 * i + 1] = m->clusters[i].last;
    }
    if (bounds) bounds[i] = m->clusters[i].bounds;
  }
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// synthetic spheres
// ---------------------------------------------------------------------------
namespace {
// value noise on the unit sphere: sum of a few seeded sinusoids of the
// direction (smooth, deterministic, cheap)
struct SphereNoise {
  double k[6][3], ph[6];
  explicit SphereNoise(SplitMix64 &r) {
    for (int i = 0; i < 6; ++i) {
      for (int a = 0; a < 3; ++a) k[i][a] = r.uniform(-4.0, 4.0);
      ph[i] = r.uniform(0.0, 6.283185307179586);
    }
  }
  double operator()(const double n[3]) const {
    double s = 0.0;
    for (int i = 0; i < 6; ++i) s += std::sin(k[i][0] * n[0] + k[i][1] * n[1] + k[i][2] * n[2] + ph[i]);
    return s / 6.0;  // in [-1, 1]
  }
};
}  // namespace

extern "C" int aloha_mesh_synth_spheres(int32_t num_spheres, int32_t rings, int32_t segments,
                                        const aloha_box3f *domain, float rmin_frac,
                                        float rmax_frac, float disp_frac, uint64_t seed,
                                        aloha_mesh **out) {
  if (!out || !domain || num_spheres < 1 || rings < 2 || segments < 3 || !(rmin_frac > 0.f) ||
      rmax_frac < rmin_frac)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_synth_spheres: bad argument");
  // per sphere: 2 poles + (rings-1) x segments ring vertices,
  // 2*segments cap triangles + 2*segments*(rings-2) band triangles
  const int64_t vps = 2 + (int64_t)(rings - 1) * segments;
  const int64_t tps = 2 * (int64_t)segments * (rings - 1);
  auto *m = new aloha_mesh;
  m->verts.resize(3 * vps * num_spheres);
  m->tris.resize(3 * tps * num_spheres);
  double span[3], size = 0.0;
  for (int a = 0; a < 3; ++a) {
    span[a] = (double)domain->upper[a] - domain->lower[a];
    size = std::max(size, span[a]);
  }
  struct Sph {
    double c[3], r;
    uint64_t seed;
  };
  std::vector<Sph> sph(num_spheres);
  SplitMix64 rng(seed);
  for (auto &s : sph) {
    s.r = rng.uniform(rmin_frac, rmax_frac) * size;
    for (int a = 0; a < 3; ++a) s.c[a] = domain->lower[a] + rng.uniform(s.r, span[a] - s.r);
    s.seed = rng.next();
  }
  const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int64_t i = t; i < num_spheres; i += nt) {
        const Sph &s = sph[i];
        SplitMix64 r(s.seed);
        SphereNoise noise(r);
        float *v = &m->verts[3 * vps * i];
        auto put = [&](int64_t k, double th_, double ph_) {
          const double n[3] = {std::sin(th_) * std::cos(ph_), std::cos(th_), std::sin(th_) * std::sin(ph_)};
          const double rad = s.r * (1.0 + disp_frac * noise(n));
          for (int a = 0; a < 3; ++a) v[3 * k + a] = (float)(s.c[a] + rad * n[a]);
        };
        put(0, 0.0, 0.0);
        put(1, M_PI, 0.0);
        for (int ri = 1; ri < rings; ++ri)
          for (int si = 0; si < segments; ++si)
            put(2 + (int64_t)(ri - 1) * segments + si, M_PI * ri / rings, 2.0 * M_PI * si / segments);
        int32_t *tr = &m->tris[3 * tps * i];
        const int32_t base = (int32_t)(vps * i);
        auto ring = [&](int ri, int si) { return base + 2 + (ri - 1) * segments + (si % segments); };
        int64_t k = 0;
        for (int si = 0; si < segments; ++si) {
          tr[3 * k] = base, tr[3 * k + 1] = ring(1, si + 1), tr[3 * k + 2] = ring(1, si), ++k;
          tr[3 * k] = base + 1, tr[3 * k + 1] = ring(rings - 1, si), tr[3 * k + 2] = ring(rings - 1, si + 1), ++k;
        }
        for (int ri = 1; ri < rings - 1; ++ri)
          for (int si = 0; si < segments; ++si) {
            const int32_t a0 = ring(ri, si), a1 = ring(ri, si + 1), b0 = ring(ri + 1, si), b1 = ring(ri + 1, si + 1);
            tr[3 * k] = a0, tr[3 * k + 1] = a1, tr[3 * k + 2] = b1, ++k;
            tr[3 * k] = a0, tr[3 * k + 1] = b1, tr[3 * k + 2] = b0, ++k;
          }
      }
    });
  for (auto &x : th) x.join();
  box_empty(m->bounds);
  for (size_t i = 0; i < m->verts.size() / 3; ++i) box_extend(m->bounds, &m->verts[3 * i]);
  *out = m;
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// MeshSplitter, median strategy (chopSuey.cpp:60-198)
// ---------------------------------------------------------------------------
extern "C" int aloha_mesh_split(aloha_mesh *m, int32_t num_clusters, int32_t *count) {
  if (!m || num_clusters < 1) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_split: bad argument");
  const int64_t nt = (int64_t)m->tris.size() / 3;
  if (nt < 1) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_split: empty mesh");
  struct PrimRef {
    uint32_t prim;
    float c[3];
  };
  // primRefs from the original order: centroid of each triangle's bounds
  std::vector<PrimRef> refs(nt);
  for (int64_t i = 0; i < nt; ++i) {
    aloha_box3f b;
    box_empty(b);
    for (int k = 0; k < 3; ++k) box_extend(b, &m->verts[3 * m->tris[3 * i + k]]);
    refs[i].prim = (uint32_t)i;
    for (int a = 0; a < 3; ++a) refs[i].c[a] = (b.lower[a] + b.upper[a]) * .5f;
  }
  m->clusters.assign(1, MeshCluster{0, nt, m->bounds});
  auto minv = [&](int64_t t, int axis) {
    const int32_t *ix = &m->tris[3 * t];
    return fminf(m->verts[3 * ix[0] + axis], fminf(m->verts[3 * ix[1] + axis], m->verts[3 * ix[2] + axis]));
  };
  std::vector<int32_t> tmp;
  do {
    // the cluster with the most primitives (strict >, first wins)
    size_t pick = 0;
    int64_t maxn = 0;
    for (size_t i = 0; i < m->clusters.size(); ++i)
      if (m->clusters[i].last - m->clusters[i].first > maxn) {
        pick = i;
        maxn = m->clusters[i].last - m->clusters[i].first;
      }
    const MeshCluster dom = m->clusters[pick];
    m->clusters.erase(m->clusters.begin() + (long)pick);
    float s[3];
    for (int a = 0; a < 3; ++a) s[a] = dom.bounds.upper[a] - dom.bounds.lower[a];
    int axis = 0;
    if (s[1] > s[0] && s[1] >= s[2]) axis = 1;
    else if (s[2] > s[0] && s[2] >= s[1]) axis = 2;
    std::sort(refs.begin() + dom.first, refs.begin() + dom.last,
              [axis](const PrimRef &a, const PrimRef &b) {
                return a.c[axis] < b.c[axis] || (a.c[axis] == b.c[axis] && a.prim < b.prim);
              });
    // num counts all primitives, not the cluster's (chopSuey.cpp:146)
    const int64_t num = (int64_t)(size_t)((float)refs.size() / (float)num_clusters);
    const int64_t at = std::min<int64_t>(dom.first + num, (int64_t)refs.size() - 1);
    const float plane = refs[at].c[axis];
    // stable partition of the cluster's triangles by min vertex < plane
    tmp.clear();
    int64_t w = dom.first;
    for (int64_t t = dom.first; t < dom.last; ++t) {
      if (minv(t, axis) < plane) {
        for (int k = 0; k < 3; ++k) m->tris[3 * w + k] = m->tris[3 * t + k];
        ++w;
      } else {
        tmp.insert(tmp.end(), &m->tris[3 * t], &m->tris[3 * t + 3]);
      }
    }
    std::memcpy(&m->tris[3 * w], tmp.data(), tmp.size() * sizeof(int32_t));
    // first primitive with min vertex >= plane; domain.first if there is none
    const int64_t split = w < dom.last ? w : dom.first;
    MeshCluster L{dom.first, split, dom.bounds}, R{split, dom.last, dom.bounds};
    L.bounds.upper[axis] = plane;
    R.bounds.lower[axis] = plane;
    const size_t before = m->clusters.size();
    if (L.last - L.first > 0) m->clusters.push_back(L);
    if (R.last - R.first > 0) m->clusters.push_back(R);
    // One side empty: the cluster returns with its domain clipped at the
    // plane and the loop goes on, as the reference's does.  When the clip
    // changed nothing, the next pick repeats this step exactly: the
    // reference recurses without end there, so this stops with an error.
    if (m->clusters.size() < before + 2 &&
        std::memcmp(&m->clusters.back().bounds, &dom.bounds, sizeof(aloha_box3f)) == 0) {
      m->clusters.clear();
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_split: a cluster cannot be split");
    }
  } while ((int)m->clusters.size() < num_clusters);
  if (count) *count = (int32_t)m->clusters.size();
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// .tri files
// ---------------------------------------------------------------------------
extern "C" int aloha_tri_write(const char *path, const aloha_mesh *m) {
  if (!path || !m) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_tri_write: NULL");
  std::ofstream f(path, std::ios::binary);
  if (!f) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_write: cannot open %s", path);
  std::vector<MeshCluster> cl = m->clusters;
  if (cl.empty()) cl.push_back(MeshCluster{0, (int64_t)m->tris.size() / 3, m->bounds});
  const uint64_t nc = cl.size(), nv = m->verts.size() / 3;
  f.write((const char *)&nc, 8);
  f.write((const char *)&m->bounds, sizeof(aloha_box3f));
  f.write((const char *)&nv, 8);
  f.write((const char *)m->verts.data(), (std::streamsize)(m->verts.size() * 4));
  for (const MeshCluster &c : cl) {
    const uint64_t ni = (uint64_t)(c.last - c.first);
    f.write((const char *)&ni, 8);
    f.write((const char *)&c.bounds, sizeof(aloha_box3f));
    f.write((const char *)&m->tris[3 * c.first], (std::streamsize)(ni * 12));
  }
  if (!f) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_write: write failed");
  return ALOHA_OK;
}

extern "C" int aloha_tri_read(const char *path, aloha_mesh **out) {
  if (!path || !out) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_tri_read: NULL");
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_read: cannot open %s", path);
  const uint64_t size = (uint64_t)f.tellg();
  f.seekg(0);
  uint64_t nc = 0, nv = 0;
  auto *m = new aloha_mesh;
  f.read((char *)&nc, 8);
  f.read((char *)&m->bounds, sizeof(aloha_box3f));
  f.read((char *)&nv, 8);
  if (!f || nv > size / 12 || nc > (1u << 24)) {
    delete m;
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_read: truncated header in %s", path);
  }
  m->verts.resize(3 * nv);
  f.read((char *)m->verts.data(), (std::streamsize)(nv * 12));
  for (uint64_t i = 0; i < nc; ++i) {
    uint64_t ni = 0;
    MeshCluster c;
    f.read((char *)&ni, 8);
    f.read((char *)&c.bounds, sizeof(aloha_box3f));
    if (!f || ni > size / 12) {
      delete m;
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_read: truncated cluster %llu", (unsigned long long)i);
    }
    c.first = (int64_t)m->tris.size() / 3;
    c.last = c.first + (int64_t)ni;
    m->tris.resize(3 * c.last);
    f.read((char *)&m->tris[3 * c.first], (std::streamsize)(ni * 12));
    if (!f) {
      delete m;
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_read: truncated indices");
    }
    m->clusters.push_back(c);
  }
  for (int32_t ix : m->tris)
    if (ix < 0 || (uint64_t)ix >= nv) {
      delete m;
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_tri_read: vertex index out of range");
    }
  *out = m;
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// binned-SAH BVH
// ---------------------------------------------------------------------------
namespace aloha {

namespace {

double half_area(const aloha_box3f &b) {
  const double dx = std::max(0.f, b.upper[0] - b.lower[0]);
  const double dy = std::max(0.f, b.upper[1] - b.lower[1]);
  const double dz = std::max(0.f, b.upper[2] - b.lower[2]);
  return dx * dy + dy * dz + dz * dx;
}

// levels of halving until n triangles fit one leaf
int need_levels(int64_t n) {
  int k = 0;
  for (int64_t cap = kMaxLeaf; cap < n; cap *= 2) ++k;
  return k;
}

// Node of the binary build: both children's boxes and refs.
struct BinNode {
  float lo[2][3], hi[2][3];
  int32_t ref[2];
};

// One child of a 4-wide node while it is being collapsed: a binary-build ref
// (node index >= 0 or leaf) and its padded box.
struct Kid {
  int32_t ref;
  float lo[3], hi[3];
};

// Collapse the binary subtree under `b` into 4-wide node `at` (already
// allocated): the children are b's two, and while fewer than four, the inner
// child with the largest box (half area) is replaced by its own two; then
// each inner child becomes a 4-wide node in turn (depth-first after `at`).
void collapse(const std::vector<BinNode> &bin, int32_t b, int32_t at, int depth, TriBvh &out) {
  std::vector<Kid> kids;
  for (int s = 0; s < 2; ++s) {
    Kid k{bin[b].ref[s], {}, {}};
    for (int a = 0; a < 3; ++a) k.lo[a] = bin[b].lo[s][a], k.hi[a] = bin[b].hi[s][a];
    kids.push_back(k);
  }
  while (kids.size() < 4) {
    int pick = -1;
    double best = -1.0;
    for (int i = 0; i < (int)kids.size(); ++i) {
      if (kids[i].ref < 0) continue;
      aloha_box3f bb;
      for (int a = 0; a < 3; ++a) bb.lower[a] = kids[i].lo[a], bb.upper[a] = kids[i].hi[a];
      const double ha = half_area(bb);
      if (ha > best) best = ha, pick = i;
    }
    if (pick < 0) break;
    const int32_t c = kids[pick].ref;
    kids.erase(kids.begin() + pick);
    for (int s = 0; s < 2; ++s) {
      Kid k{bin[c].ref[s], {}, {}};
      for (int a = 0; a < 3; ++a) k.lo[a] = bin[c].lo[s][a], k.hi[a] = bin[c].hi[s][a];
      kids.insert(kids.begin() + pick + s, k);
    }
  }
  int32_t refs[4];
  for (int i = 0; i < 4; ++i) {
    TriNode4 &n = out.nodes[at];
    if (i >= (int)kids.size()) {
      n.lo_x[i] = n.lo_y[i] = n.lo_z[i] = FLT_MAX;
      n.hi_x[i] = n.hi_y[i] = n.hi_z[i] = -FLT_MAX;
      refs[i] = kEmptyChild;
      continue;
    }
    const Kid &k = kids[i];
    n.lo_x[i] = k.lo[0], n.hi_x[i] = k.hi[0];
    n.lo_y[i] = k.lo[1], n.hi_y[i] = k.hi[1];
    n.lo_z[i] = k.lo[2], n.hi_z[i] = k.hi[2];
    if (k.ref < 0) {
      refs[i] = k.ref;
      out.depth = std::max(out.depth, depth + 1);
    } else {
      refs[i] = (int32_t)out.nodes.size();
      out.nodes.push_back(TriNode4{});
      collapse(bin, k.ref, refs[i], depth + 1, out);
    }
  }
  for (int i = 0; i < 4; ++i) out.nodes[at].ref[i] = refs[i];
}

}  // namespace

bool build_tri_bvh(const float *tri9, int64_t ntris, float pad, TriBvh &out, std::string &err) {
  out = TriBvh{};
  out.pad = pad;
  if (ntris <= 0) {
    err = "empty geometry";
    return false;
  }
  if (ntris >= (int64_t(1) << (31 - kLeafBits))) {
    err = "too many triangles in one geometry for the leaf encoding";
    return false;
  }
  // 8 centroid bins per axis: on the 4096-sphere scene (1920x1080 primary
  // rays, 8 ranks) the walk takes 1.66 ms with 8 bins against 2.25 with 16
  // and 3.4 with 12 or 32 -- finer bins let the SAH peel thin slabs off
  // clusters of overlapping spheres, which its uniform-ray cost model rates
  // well and the camera's rays do not (same images in every case)
  constexpr int kBins = 8;
  constexpr double kTrav = 1.0;  // cost of one node step relative to one triangle test
  std::vector<uint32_t> &order = out.order;
  order.resize(ntris);
  std::iota(order.begin(), order.end(), 0u);
  std::vector<float> cen(3 * ntris);
  std::vector<aloha_box3f> tb(ntris);
  for (int64_t i = 0; i < ntris; ++i) {
    box_empty(tb[i]);
    for (int k = 0; k < 3; ++k) box_extend(tb[i], tri9 + 9 * i + 3 * k);
    for (int a = 0; a < 3; ++a) cen[3 * i + a] = (tb[i].lower[a] + tb[i].upper[a]) * .5f;
  }
  struct Task {
    int32_t parent, side, depth;
    int64_t first, last;
  };
  std::vector<BinNode> bin;  // the binary build, collapsed below
  int32_t broot = 0;
  std::vector<Task> stack{{-1, 0, 0, 0, ntris}};
  while (!stack.empty()) {
    const Task t = stack.back();
    stack.pop_back();
    aloha_box3f b, cb;
    box_empty(b);
    box_empty(cb);
    for (int64_t i = t.first; i < t.last; ++i) {
      const uint32_t p = order[i];
      box_extend(b, tb[p].lower);
      box_extend(b, tb[p].upper);
      box_extend(cb, &cen[3 * p]);
    }
    const int64_t n = t.last - t.first;
    // depth bound: leaves at most kMaxBvhDepth deep.  A task may use SAH as
    // long as halving from one level down still fits; at the bound it splits
    // at the object median, which lowers the levels it needs by one.
    const int need = need_levels(n);
    const bool must_median = n > kMaxLeaf && t.depth + need >= kMaxBvhDepth;
    int axis = -1, split_bin = 0;
    bool median = must_median;
    if (!must_median && n > 1 && t.depth < kMaxBvhDepth) {
      double best = n <= kMaxLeaf ? (double)n : 1e300;  // leaf only when small enough
      const double pa = std::max(half_area(b), 1e-30);
      for (int a = 0; a < 3; ++a) {
        const float ext = cb.upper[a] - cb.lower[a];
        if (!(ext > 0.f)) continue;
        const float scale = kBins / ext;
        aloha_box3f bb[kBins];
        int64_t cnt[kBins] = {};
        for (auto &x : bb) box_empty(x);
        for (int64_t i = t.first; i < t.last; ++i) {
          const uint32_t p = order[i];
          const int k = std::min(kBins - 1, std::max(0, (int)((cen[3 * p + a] - cb.lower[a]) * scale)));
          ++cnt[k];
          box_extend(bb[k], tb[p].lower);
          box_extend(bb[k], tb[p].upper);
        }
        double rA[kBins];
        int64_t rN[kBins];
        aloha_box3f acc;
        box_empty(acc);
        int64_t c = 0;
        for (int k = kBins - 1; k >= 0; --k) {
          box_extend(acc, bb[k].lower);
          box_extend(acc, bb[k].upper);
          c += cnt[k];
          rA[k] = c ? half_area(acc) : 0.0;
          rN[k] = c;
        }
        box_empty(acc);
        c = 0;
        for (int k = 0; k < kBins - 1; ++k) {
          box_extend(acc, bb[k].lower);
          box_extend(acc, bb[k].upper);
          c += cnt[k];
          if (!c || !rN[k + 1]) continue;
          const double cost = kTrav + (half_area(acc) * (double)c + rA[k + 1] * (double)rN[k + 1]) / pa;
          if (cost < best) {
            best = cost;
            axis = a;
            split_bin = k;
          }
        }
      }
      if (axis < 0 && n > kMaxLeaf) median = true;  // no usable plane (coincident centroids)
    }
    int32_t ref;
    int64_t mid = 0;
    if (axis < 0 && !median) {
      ref = ~(int32_t)((t.first << kLeafBits) | (n - 1));
      out.depth2 = std::max(out.depth2, t.depth);
    } else {
      if (median) {
        int a = 0;
        for (int k = 1; k < 3; ++k)
          if (cb.upper[k] - cb.lower[k] > cb.upper[a] - cb.lower[a]) a = k;
        mid = t.first + n / 2;
        std::nth_element(order.begin() + t.first, order.begin() + mid, order.begin() + t.last,
                         [&](uint32_t x, uint32_t y) {
                           return cen[3 * x + a] < cen[3 * y + a] || (cen[3 * x + a] == cen[3 * y + a] && x < y);
                         });
      } else {
        const float ext = cb.upper[axis] - cb.lower[axis], scale = kBins / ext;
        const float lo = cb.lower[axis];
        auto it = std::partition(order.begin() + t.first, order.begin() + t.last, [&](uint32_t p) {
          const int k = std::min(kBins - 1, std::max(0, (int)((cen[3 * p + axis] - lo) * scale)));
          return k <= split_bin;
        });
        mid = it - order.begin();
      }
      ref = (int32_t)bin.size();
      bin.push_back(BinNode{});
      // right first: the left subtree is built next, right after its parent
      stack.push_back({ref, 1, t.depth + 1, mid, t.last});
      stack.push_back({ref, 0, t.depth + 1, t.first, mid});
    }
    if (t.parent < 0) {
      broot = ref;
      for (int a = 0; a < 3; ++a) {
        out.lo[a] = b.lower[a];
        out.hi[a] = b.upper[a];
      }
      continue;
    }
    BinNode &pn = bin[t.parent];
    for (int a = 0; a < 3; ++a) {
      pn.lo[t.side][a] = b.lower[a] - pad;
      pn.hi[t.side][a] = b.upper[a] + pad;
    }
    pn.ref[t.side] = ref;
  }
  // the 4-wide tree: a leaf root stays the root; else node 0 collapses the
  // binary root
  if (broot < 0) {
    out.root = broot;
  } else {
    out.root = 0;
    out.nodes.reserve(bin.size() / 2 + 1);
    out.nodes.push_back(TriNode4{});
    collapse(bin, broot, 0, 0, out);
  }
  return true;
}

void flatten(const aloha_mesh *m, const int32_t *clusters, const uint32_t *colors, int n,
             RankBuild &rb) {
  int64_t total = 0;
  for (int k = 0; k < n; ++k) {
    const MeshCluster &c = m->clusters[clusters[k]];
    total += c.last - c.first;
  }
  rb.tri9.resize(9 * total);
  rb.color.resize(total);
  int64_t w = 0;
  for (int k = 0; k < n; ++k) {
    const MeshCluster &c = m->clusters[clusters[k]];
    for (int64_t t = c.first; t < c.last; ++t, ++w) {
      for (int v = 0; v < 3; ++v)
        std::memcpy(&rb.tri9[9 * w + 3 * v], &m->verts[3 * (size_t)m->tris[3 * t + v]], 12);
      rb.color[w] = colors[k];
    }
  }
}

void build_rank(RankBuild &rb) {
  const int64_t nt = (int64_t)rb.color.size();
  if (nt == 0) return;
  // box padding: 2^-18 of the geometry's coordinate scale, far above the
  // rounding of the slab test at any camera within ~100x that scale
  float scale = 0.f;
  for (float v : rb.tri9) scale = std::max(scale, std::fabs(v));
  const float pad = std::ldexp(std::max(scale, 1e-30f), -18);
  if (!build_tri_bvh(rb.tri9.data(), nt, pad, rb.bvh, rb.err)) {
    rb.ok = false;
    return;
  }
  rb.dev_tris.resize(12 * (size_t)nt);
  for (int64_t s = 0; s < nt; ++s) {
    const uint32_t p = rb.bvh.order[s];
    float *o = &rb.dev_tris[12 * s];
    for (int v = 0; v < 3; ++v) std::memcpy(o + 4 * v, &rb.tri9[9 * (size_t)p + 3 * v], 12);
    const int32_t prim = (int32_t)p;
    std::memcpy(o + 3, &prim, 4);
    std::memcpy(o + 7, &rb.color[p], 4);
    std::memset(o + 11, 0, 4);
  }
}

}  // namespace aloha

// ---------------------------------------------------------------------------
// host-only build of one rank's device layout
// ---------------------------------------------------------------------------
extern "C" int aloha_tris_build_host(const aloha_mesh *mesh, int32_t n, const int32_t *cluster_ids,
                                     const uint32_t *color_ids, int64_t *ntris, int64_t *nnodes,
                                     int32_t *depth, void *nodes, float *tris, int32_t *root) {
  if (!mesh || n < 0 || (n > 0 && (!cluster_ids || !color_ids)))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_tris_build_host: bad argument");
  aloha_mesh whole;
  const aloha_mesh *m = mesh;
  if (mesh->clusters.empty()) {
    whole.verts = mesh->verts;
    whole.tris = mesh->tris;
    whole.bounds = mesh->bounds;
    whole.clusters.push_back(MeshCluster{0, (int64_t)mesh->tris.size() / 3, mesh->bounds});
    m = &whole;
  }
  for (int k = 0; k < n; ++k)
    if (cluster_ids[k] < 0 || cluster_ids[k] >= (int)m->clusters.size())
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_tris_build_host: cluster id %d", cluster_ids[k]);
  RankBuild rb;
  flatten(m, cluster_ids, color_ids, n, rb);
  build_rank(rb);
  if (!rb.ok) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_tris_build_host: %s", rb.err.c_str());
  const int64_t nt = (int64_t)rb.color.size();
  if (ntris) *ntris = nt;
  if (nnodes) *nnodes = (int64_t)rb.bvh.nodes.size();
  if (depth) *depth = rb.bvh.depth;
  if (root) *root = nt ? rb.bvh.root : INT32_MIN;
  if (nodes && !rb.bvh.nodes.empty())
    std::memcpy(nodes, rb.bvh.nodes.data(), rb.bvh.nodes.size() * sizeof(TriNode4));
  if (tris && nt) std::memcpy(tris, rb.dev_tris.data(), rb.dev_tris.size() * sizeof(float));
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// Wavefront OBJ input (the mesh mode of chopSuey, chopSuey.cpp:400-421, and
// Mesh::loadOBJ, mesh.cpp:39-85): positions and the triangulated faces of the
// first shape, as MeshSplitter reads geoms[0] only.  Polygons are fanned
// (v0, vi, vi+1); indices may be negative (relative to the vertices read so
// far) or carry /vt/vn parts, which are ignored.
// ---------------------------------------------------------------------------
extern "C" int aloha_mesh_read_obj(const char *path, aloha_mesh **out) {
  if (!path || !out) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_mesh_read_obj: NULL");
  std::ifstream f(path);
  if (!f) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_mesh_read_obj: cannot open %s", path);
  auto *m = new aloha_mesh;
  std::string line;
  bool shape_has_faces = false, done = false;
  int64_t lineno = 0;
  std::vector<int64_t> poly;
  while (!done && std::getline(f, line)) {
    ++lineno;
    const char *s = line.c_str();
    while (*s == ' ' || *s == '\t') ++s;
    if ((s[0] == 'o' || s[0] == 'g') && (s[1] == ' ' || s[1] == '\t' || s[1] == 0)) {
      if (shape_has_faces) done = true;  // a second shape starts: geoms[0] is complete
      continue;
    }
    if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
      float v[3];
      char *end = nullptr;
      const char *p = s + 2;
      for (int k = 0; k < 3; ++k) {
        v[k] = std::strtof(p, &end);
        if (end == p) {
          delete m;
          return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_mesh_read_obj: bad vertex on line %lld",
                                   (long long)lineno);
        }
        p = end;
      }
      m->verts.insert(m->verts.end(), v, v + 3);
      continue;
    }
    if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
      poly.clear();
      const char *p = s + 2;
      const int64_t nv = (int64_t)m->verts.size() / 3;
      while (true) {
        while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
        if (!*p) break;
        char *end = nullptr;
        const long long idx = std::strtoll(p, &end, 10);
        if (end == p || idx == 0) {
          delete m;
          return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_mesh_read_obj: bad face on line %lld",
                                   (long long)lineno);
        }
        const int64_t v = idx > 0 ? idx - 1 : nv + idx;
        if (v < 0 || v >= nv) {
          delete m;
          return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_mesh_read_obj: index out of range on line %lld",
                                   (long long)lineno);
        }
        poly.push_back(v);
        p = end;
        while (*p && *p != ' ' && *p != '\t' && *p != '\r') ++p;  // skip /vt/vn
      }
      if (poly.size() < 3) {
        delete m;
        return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_mesh_read_obj: face with %d vertices on line %lld",
                                 (int)poly.size(), (long long)lineno);
      }
      for (size_t k = 1; k + 1 < poly.size(); ++k) {
        m->tris.push_back((int32_t)poly[0]);
        m->tris.push_back((int32_t)poly[k]);
        m->tris.push_back((int32_t)poly[k + 1]);
      }
      shape_has_faces = true;
    }
  }
  // vertices listed after the first shape still belong to the shared array
  while (std::getline(f, line)) {
    const char *s = line.c_str();
    while (*s == ' ' || *s == '\t') ++s;
    if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
      float v[3] = {0, 0, 0};
      char *end = nullptr;
      const char *p = s + 2;
      for (int k = 0; k < 3; ++k) {
        v[k] = std::strtof(p, &end);
        p = end;
      }
      m->verts.insert(m->verts.end(), v, v + 3);
    }
  }
  if (m->verts.size() / 3 > (size_t)INT32_MAX) {
    delete m;
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_mesh_read_obj: too many vertices");
  }
  box_empty(m->bounds);
  for (size_t i = 0; i < m->verts.size() / 3; ++i) box_extend(m->bounds, &m->verts[3 * i]);
  *out = m;
  return ALOHA_OK;
}
