// host_bricking.cpp -- brick layout, brick->rank assignment, visibility
// order and the .vol file format, all on the host.
//
//   aloha_split_volume      VolumeSplitter (chopSuey.cpp:245-323)
//   aloha_volume_extract    Volume::get (volume.h:46-98)
//   aloha_partitioner_*     Partitioner (waikikiRenderer.cpp:126-338)
//   aloha_vol_*             .vol writer/reader (chopSuey.cpp:325-349,
//                           waikikiRenderer.cpp:596-665)
//
// These decide which voxels each brick owns, which rank owns each brick and
// in which order ranks are blended, so they reproduce the reference's
// integer and float comparisons exactly (strict '>' picks, truncating plane
// conversion, centroid '<' plane).
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <numeric>
#include <vector>

#include "internal.h"

namespace aloha {

std::string &global_error() {
  static thread_local std::string s;
  return s;
}

namespace {

struct Domain {
  aloha_box3i cell, voxel;
  aloha_box3f space;
};

float space_volume(const aloha_box3f &b) {
  // owl box3f::volume() = size.x*size.y*size.z, evaluated left to right
  float sx = b.upper[0] - b.lower[0];
  float sy = b.upper[1] - b.lower[1];
  float sz = b.upper[2] - b.lower[2];
  return sx * sy * sz;
}

long long cell_volume(const aloha_box3i &b) {
  return (long long)(b.upper[0] - b.lower[0]) * (b.upper[1] - b.lower[1]) *
         (b.upper[2] - b.lower[2]);
}

// Split-axis rule shared by VolumeSplitter (integer sizes) and partitionKD
// (float sizes): x unless y is strictly largest against x and not smaller
// than z, then z under the mirrored rule.
template <typename T>
int split_axis(const T s[3]) {
  if (s[1] > s[0] && s[1] >= s[2]) return 1;
  if (s[2] > s[0] && s[2] >= s[1]) return 2;
  return 0;
}

}  // namespace
}  // namespace aloha

using namespace aloha;

extern "C" const char *aloha_version(void) { return "aloha_amd 0.1 (gfx950)"; }

extern "C" int aloha_split_volume(const int32_t dims[3], int32_t num_bricks,
                                  aloha_brick_desc *out, int32_t cap, int32_t *count) {
  if (!dims || !out || !count || cap < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_split_volume: null argument or cap < 1");
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 1)
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_split_volume: dims must be >= 1");
  std::vector<Domain> cl;
  Domain d{};
  for (int a = 0; a < 3; ++a) {
    d.cell.upper[a] = dims[a];
    d.voxel.upper[a] = dims[a] + 1;  // header voxelRange = [0, dims+1) (chopSuey.cpp:251)
    d.space.upper[a] = (float)dims[a];
  }
  cl.push_back(d);
  // doSplit runs at least once, then until num_bricks bricks exist.
  do {
    size_t pick = 0;
    float maxv = 0.f;
    for (size_t i = 0; i < cl.size(); ++i) {
      float v = space_volume(cl[i].space);
      if (v > maxv) {
        pick = i;
        maxv = v;
      }
    }
    Domain dom = cl[pick];
    cl.erase(cl.begin() + (long)pick);
    int sz[3];
    for (int a = 0; a < 3; ++a) sz[a] = dom.cell.upper[a] - dom.cell.lower[a];
    const int axis = split_axis(sz);
    const int plane = (int)((float)dom.cell.lower[axis] + (float)sz[axis] * .5f);
    Domain L = dom, R = dom;
    L.cell.upper[axis] = plane;
    R.cell.lower[axis] = plane;
    for (Domain *c : {&L, &R}) {
      for (int a = 0; a < 3; ++a) {
        c->voxel.lower[a] = c->cell.lower[a];
        c->voxel.upper[a] = std::min(c->cell.upper[a] + 1, dims[a]);
        c->space.lower[a] = (float)c->cell.lower[a];
        c->space.upper[a] = (float)c->cell.upper[a];
      }
    }
    const size_t before = cl.size();
    if (cell_volume(L.cell) > 0) cl.push_back(L);
    if (cell_volume(R.cell) > 0) cl.push_back(R);
    // The reference recurses forever when a split yields a single non-empty
    // half (a one-cell brick); report it instead.
    if (cl.size() < before + 2)
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_split_volume: cannot split a one-cell brick");
  } while ((int)cl.size() < num_bricks);
  if ((int)cl.size() > cap)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_split_volume: %zu bricks exceed cap %d",
                             cl.size(), cap);
  for (size_t i = 0; i < cl.size(); ++i) {
    out[i].cell_range = cl[i].cell;
    out[i].voxel_range = cl[i].voxel;
    out[i].space_range = cl[i].space;
    out[i].value_range = {FLT_MAX, -FLT_MAX};
  }
  *count = (int32_t)cl.size();
  return ALOHA_OK;
}

extern "C" int aloha_volume_extract(const void *raw, const int32_t dims[3], int32_t bpc,
                                    const aloha_box3i *range, float *out,
                                    aloha_range1f *value_range) {
  if (!raw || !dims || !range || !out)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_volume_extract: null argument");
  if (bpc != 1 && bpc != 2 && bpc != 4)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_volume_extract: bpc must be 1, 2 or 4");
  for (int a = 0; a < 3; ++a)
    if (range->lower[a] < 0 || range->upper[a] > dims[a] || range->lower[a] > range->upper[a])
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_volume_extract: range outside volume");
  float lo = FLT_MAX, hi = -FLT_MAX;
  size_t o = 0;
  const int nx = range->upper[0] - range->lower[0];
  for (int z = range->lower[2]; z < range->upper[2]; ++z)
    for (int y = range->lower[1]; y < range->upper[1]; ++y) {
      const size_t first = (size_t)z * dims[0] * (size_t)dims[1] + (size_t)y * dims[0] +
                           (size_t)range->lower[0];
      for (int ix = 0; ix < nx; ++ix) {
        float v;
        if (bpc == 1) v = (float)((const uint8_t *)raw)[first + ix] / 255.f;
        else if (bpc == 2) v = (float)((const uint16_t *)raw)[first + ix] / 65535.f;
        else v = ((const float *)raw)[first + ix];
        out[o++] = v;
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
    }
  if (value_range) {
    value_range->lower = std::min(value_range->lower, lo);
    value_range->upper = std::max(value_range->upper, hi);
  }
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// Partitioner
// ---------------------------------------------------------------------------
struct KDNode {
  int32_t axis;
  float plane;
  int32_t child1, child2;
};
static_assert(sizeof(KDNode) == 16, "KDNode layout");

struct aloha_partitioner {
  int num_ranks = 0;
  std::vector<std::vector<int32_t>> per_rank;
  std::vector<KDNode> kd;
};

extern "C" int aloha_partitioner_create_kd(const aloha_box3f *domains, int32_t n,
                                           int32_t num_ranks, aloha_partitioner **out) {
  if (!out || (n > 0 && !domains) || n < 0 || num_ranks < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_partitioner_create_kd: bad argument");
  auto *p = new aloha_partitioner;
  p->num_ranks = num_ranks;
  p->per_rank.resize(num_ranks);
  const int per_rank = n / num_ranks;

  struct Set {
    aloha_box3f bounds;
    std::vector<int32_t> ids;
    int kd;
  };
  auto empty_box = [] {
    aloha_box3f b;
    for (int a = 0; a < 3; ++a) {
      b.lower[a] = 1e30f;
      b.upper[a] = -1e30f;
    }
    return b;
  };
  auto extend = [](aloha_box3f &b, const aloha_box3f &d) {
    for (int a = 0; a < 3; ++a) {
      b.lower[a] = std::min(b.lower[a], d.lower[a]);
      b.upper[a] = std::max(b.upper[a], d.upper[a]);
    }
  };
  Set all{empty_box(), {}, 0};
  for (int i = 0; i < n; ++i) {
    extend(all.bounds, domains[i]);
    all.ids.push_back(i);
  }
  std::vector<Set> open{all}, assigned;
  p->kd.push_back(KDNode{0, 0.f, INT_MAX, INT_MAX});
  while (!open.empty()) {
    size_t pick = 0, maxn = 0;
    for (size_t i = 0; i < open.size(); ++i)
      if (open[i].ids.size() > maxn) {
        pick = i;
        maxn = open[i].ids.size();
      }
    Set cs = open[pick];
    open.erase(open.begin() + (long)pick);
    if ((int)cs.ids.size() <= per_rank || (int)assigned.size() == num_ranks - 1) {
      const int rank = (int)assigned.size();
      p->kd[cs.kd].child1 = ~rank;
      p->kd[cs.kd].child2 = ~rank;
      assigned.push_back(cs);
      if ((int)assigned.size() == num_ranks) {
        // the last rank takes every set still open (waikikiRenderer.cpp:200-211)
        Set &last = assigned.back();
        for (const Set &o : open) last.ids.insert(last.ids.end(), o.ids.begin(), o.ids.end());
        open.clear();
        break;
      }
      continue;
    }
    float s[3];
    for (int a = 0; a < 3; ++a) s[a] = cs.bounds.upper[a] - cs.bounds.lower[a];
    const int axis = split_axis(s);
    const float plane = cs.bounds.lower[axis] + s[axis] * .5f;
    Set L{empty_box(), {}, 0}, R{empty_box(), {}, 0};
    for (int32_t id : cs.ids) {
      const aloha_box3f &d = domains[id];
      const float c = (d.lower[axis] + d.upper[axis]) * .5f;  // box3f::center()
      if (c < plane) {
        extend(L.bounds, d);
        L.ids.push_back(id);
      } else {
        extend(R.bounds, d);
        R.ids.push_back(id);
      }
    }
    // The reference picks the same set again forever when the middle plane
    // separates nothing (e.g. fewer bricks than ranks); report it instead.
    if (L.ids.empty() || R.ids.empty()) {
      delete p;
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG,
                               "aloha_partitioner_create_kd: %zu bricks cannot be separated "
                               "for %d ranks",
                               cs.ids.size(), num_ranks);
    }
    p->kd[cs.kd].axis = axis;
    p->kd[cs.kd].plane = plane;
    L.kd = (int)p->kd.size();
    p->kd[cs.kd].child1 = L.kd;
    p->kd.push_back(KDNode{0, 0.f, INT_MAX, INT_MAX});
    open.push_back(std::move(L));
    R.kd = (int)p->kd.size();
    p->kd[cs.kd].child2 = R.kd;
    p->kd.push_back(KDNode{0, 0.f, INT_MAX, INT_MAX});
    open.push_back(std::move(R));
  }
  for (size_t r = 0; r < assigned.size(); ++r) p->per_rank[r] = assigned[r].ids;
  *out = p;
  return ALOHA_OK;
}

extern "C" int aloha_partitioner_create_rr(int32_t n, int32_t num_ranks,
                                           aloha_partitioner **out) {
  if (!out || n < 0 || num_ranks < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_partitioner_create_rr: bad argument");
  auto *p = new aloha_partitioner;
  p->num_ranks = num_ranks;
  p->per_rank.resize(num_ranks);
  const int per = (n + num_ranks - 1) / num_ranks;
  for (int i = 0; i < n; ++i) {
    const int r = i / per;
    if (r < num_ranks) p->per_rank[r].push_back(i);
  }
  *out = p;
  return ALOHA_OK;
}

extern "C" void aloha_partitioner_destroy(aloha_partitioner *p) { delete p; }

extern "C" int aloha_partitioner_rank_bricks(const aloha_partitioner *p, int32_t rank,
                                             int32_t *ids, int32_t cap, int32_t *count) {
  if (!p || !count || rank < 0 || rank >= p->num_ranks)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_partitioner_rank_bricks: bad argument");
  const auto &v = p->per_rank[rank];
  *count = (int32_t)v.size();
  if ((int)v.size() > cap || (!ids && !v.empty()))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_partitioner_rank_bricks: cap too small");
  std::copy(v.begin(), v.end(), ids);
  return ALOHA_OK;
}

extern "C" int aloha_partitioner_composite_order(const aloha_partitioner *p,
                                                 const float reference[3], int32_t *order,
                                                 int32_t cap, int32_t *count) {
  if (!p || !reference || !count)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_partitioner_composite_order: bad argument");
  std::vector<int32_t> res;
  if (p->kd.empty()) {
    // no kd tree (round-robin): identity order (waikikiRenderer.cpp:132-133, :283-286)
    res.resize(p->num_ranks);
    std::iota(res.begin(), res.end(), 0);
  } else {
    std::vector<int> stack{0};
    int addr = 0;
    while (!stack.empty()) {
      const KDNode &node = p->kd[addr];
      if (node.child1 < 0 && node.child2 < 0) {
        res.push_back(~node.child1);
        addr = stack.back();
        stack.pop_back();
      } else if (node.child1 == INT_MAX) {
        addr = stack.back();
        stack.pop_back();
      } else if (reference[node.axis] < node.plane) {
        addr = node.child1;
        stack.push_back(node.child2);
      } else {
        addr = node.child2;
        stack.push_back(node.child1);
      }
    }
  }
  *count = (int32_t)res.size();
  if ((int)res.size() > cap || !order)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_partitioner_composite_order: cap too small");
  std::copy(res.begin(), res.end(), order);
  return ALOHA_OK;
}

extern "C" int aloha_partitioner_kd_nodes(const aloha_partitioner *p, void *nodes, int32_t cap,
                                          int32_t *count) {
  if (!p || !count) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_partitioner_kd_nodes: null");
  *count = (int32_t)p->kd.size();
  if ((int)p->kd.size() > cap || (!nodes && !p->kd.empty()))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_partitioner_kd_nodes: cap too small");
  if (!p->kd.empty()) std::memcpy(nodes, p->kd.data(), p->kd.size() * sizeof(KDNode));
  return ALOHA_OK;
}

// ---------------------------------------------------------------------------
// .vol files
// ---------------------------------------------------------------------------
static_assert(sizeof(aloha_brick_desc) == 80, "brick record layout");

static size_t desc_voxels(const aloha_brick_desc &d) {
  size_t n = 1;
  for (int a = 0; a < 3; ++a) {
    const long long e = (long long)d.voxel_range.upper[a] - d.voxel_range.lower[a];
    if (e <= 0) return 0;
    n *= (size_t)e;
  }
  return n;
}

extern "C" int aloha_vol_write(const char *path, const aloha_box3i *cell_range,
                               const aloha_box3i *voxel_range, const aloha_box3f *space_range,
                               int32_t n, const aloha_brick_desc *bricks,
                               const float *const *voxels) {
  if (!path || !cell_range || !voxel_range || !space_range || n < 0 || (n > 0 && (!bricks || !voxels)))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_vol_write: bad argument");
  std::ofstream ofs(path, std::ios::binary);
  if (!ofs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_write: cannot open %s", path);
  const uint64_t nb = (uint64_t)n;
  ofs.write((const char *)&nb, sizeof nb);
  ofs.write((const char *)cell_range, sizeof *cell_range);
  ofs.write((const char *)voxel_range, sizeof *voxel_range);
  ofs.write((const char *)space_range, sizeof *space_range);
  for (int i = 0; i < n; ++i) {
    ofs.write((const char *)&bricks[i], sizeof(aloha_brick_desc));
    ofs.write((const char *)voxels[i], (std::streamsize)(sizeof(float) * desc_voxels(bricks[i])));
  }
  if (!ofs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_write: write failed on %s", path);
  return ALOHA_OK;
}

extern "C" int aloha_vol_read_header(const char *path, aloha_box3i *cell_range,
                                     aloha_box3i *voxel_range, aloha_box3f *space_range,
                                     aloha_brick_desc *bricks, int32_t cap, int32_t *n) {
  if (!path || !n) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_vol_read_header: null");
  std::ifstream ifs(path, std::ios::binary | std::ios::ate);
  if (!ifs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_header: cannot open %s", path);
  const uint64_t file_size = (uint64_t)ifs.tellg();
  ifs.seekg(0);
  uint64_t nb = 0;
  aloha_box3i cr, vr;
  aloha_box3f sr;
  ifs.read((char *)&nb, sizeof nb);
  ifs.read((char *)&cr, sizeof cr);
  ifs.read((char *)&vr, sizeof vr);
  ifs.read((char *)&sr, sizeof sr);
  if (!ifs || nb > (1u << 24))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_header: truncated header in %s", path);
  uint64_t pos = 8 + 24 + 24 + 24;
  for (uint64_t i = 0; i < nb; ++i) {
    aloha_brick_desc d;
    ifs.seekg((std::streamoff)pos);
    ifs.read((char *)&d, sizeof d);
    if (!ifs)
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_header: truncated brick %llu",
                               (unsigned long long)i);
    const size_t nv = desc_voxels(d);
    pos += sizeof d + sizeof(float) * nv;
    if (pos > file_size)
      return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_header: brick %llu voxels truncated",
                               (unsigned long long)i);
    if ((int64_t)i < cap && bricks) bricks[i] = d;
  }
  if (cell_range) *cell_range = cr;
  if (voxel_range) *voxel_range = vr;
  if (space_range) *space_range = sr;
  *n = (int32_t)nb;
  if ((int64_t)nb > cap && bricks)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_vol_read_header: cap too small");
  return ALOHA_OK;
}

extern "C" int aloha_vol_read_brick(const char *path, int32_t i, float *out, size_t out_count) {
  if (!path || !out || i < 0) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_vol_read_brick: bad argument");
  std::ifstream ifs(path, std::ios::binary);
  if (!ifs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_brick: cannot open %s", path);
  uint64_t nb = 0;
  ifs.read((char *)&nb, sizeof nb);
  if (!ifs || (uint64_t)i >= nb)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_vol_read_brick: brick %d not in file", i);
  uint64_t pos = 8 + 72;
  for (int32_t k = 0;; ++k) {
    aloha_brick_desc d;
    ifs.seekg((std::streamoff)pos);
    ifs.read((char *)&d, sizeof d);
    if (!ifs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_brick: truncated file");
    const size_t nv = desc_voxels(d);
    if (k == i) {
      if (out_count < nv)
        return ALOHA_FAIL_GLOBAL(ALOHA_ERR_LIMIT, "aloha_vol_read_brick: need %zu voxels", nv);
      ifs.read((char *)out, (std::streamsize)(sizeof(float) * nv));
      if (!ifs) return ALOHA_FAIL_GLOBAL(ALOHA_ERR_IO, "aloha_vol_read_brick: truncated voxels");
      return ALOHA_OK;
    }
    pos += sizeof d + sizeof(float) * nv;
  }
}
