// host_check.cpp -- drives the library's host-only C functions under
// AddressSanitizer and UndefinedBehaviorSanitizer: bricking and
// partitioning, .vol / .tri / .obj files, seeded synthetic inputs, the
// camera projection of boxes, the mesh split and the BVH build, including
// their error paths.  Built from the library's host sources with g++ by
// tests/test_sanitize.py (no HIP code is compiled).  argv[1]: a scratch
// directory; the BVH of one rank is written there for oracle_check.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "aloha_amd.h"

// The context library (HIP) is not built here: the message of a failing
// host call is the thread's global error, which this stands in for.
namespace aloha {
std::string &global_error();
}
extern "C" const char *aloha_last_error(const aloha_ctx *) { return aloha::global_error().c_str(); }

#define CHECK(c)                                                                  \
  do {                                                                            \
    if (!(c)) {                                                                   \
      std::fprintf(stderr, "check failed at %s:%d: %s (%s)\n", __FILE__, __LINE__, \
                   #c, aloha_last_error(nullptr));                                \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

static int64_t voxels_of(const aloha_box3i &b) {
  return (int64_t)(b.upper[0] - b.lower[0]) * (b.upper[1] - b.lower[1]) * (b.upper[2] - b.lower[2]);
}

static int bricking(const std::string &dir) {
  const int32_t dims[3] = {65, 33, 17};
  std::vector<uint8_t> raw((size_t)dims[0] * dims[1] * dims[2]);
  CHECK(aloha_synth_blobs_u8(dims, 6, 3, 2.f, 6.f, .3f, 1.f, raw.data()) == ALOHA_OK);
  aloha_brick_desc descs[16];
  int32_t nb = 0;
  CHECK(aloha_split_volume(dims, 5, descs, 16, &nb) == ALOHA_OK && nb >= 5);
  std::vector<std::vector<float>> vox(nb);
  std::vector<const float *> ptrs(nb);
  for (int i = 0; i < nb; ++i) {
    vox[i].resize((size_t)voxels_of(descs[i].voxel_range));
    CHECK(aloha_volume_extract(raw.data(), dims, 1, &descs[i].voxel_range, vox[i].data(),
                               &descs[i].value_range) == ALOHA_OK);
    ptrs[i] = vox[i].data();
  }
  // .vol round trip
  aloha_box3i cell{{0, 0, 0}, {dims[0], dims[1], dims[2]}};
  aloha_box3i vrange{{0, 0, 0}, {dims[0] + 1, dims[1] + 1, dims[2] + 1}};
  aloha_box3f space{{0, 0, 0}, {(float)dims[0], (float)dims[1], (float)dims[2]}};
  const std::string vol = dir + "/check.vol";
  CHECK(aloha_vol_write(vol.c_str(), &cell, &vrange, &space, nb, descs, ptrs.data()) == ALOHA_OK);
  aloha_brick_desc back[16];
  aloha_box3i c2, v2;
  aloha_box3f s2;
  int32_t n2 = 0;
  CHECK(aloha_vol_read_header(vol.c_str(), &c2, &v2, &s2, back, 16, &n2) == ALOHA_OK && n2 == nb);
  for (int i = 0; i < nb; ++i) {
    std::vector<float> v(vox[i].size());
    CHECK(aloha_vol_read_brick(vol.c_str(), i, v.data(), v.size()) == ALOHA_OK);
    CHECK(std::memcmp(v.data(), vox[i].data(), v.size() * sizeof(float)) == 0);
  }
  // error paths
  int32_t cnt = 0;
  CHECK(aloha_vol_read_brick(vol.c_str(), nb, vox[0].data(), vox[0].size()) != ALOHA_OK);
  CHECK(aloha_vol_read_brick(vol.c_str(), 0, vox[0].data(), 1) != ALOHA_OK);
  CHECK(aloha_vol_read_header((dir + "/missing.vol").c_str(), &c2, &v2, &s2, back, 16, &n2) != ALOHA_OK);
  CHECK(aloha_vol_read_header(vol.c_str(), &c2, &v2, &s2, back, 1, &n2) != ALOHA_OK);
  const int32_t bad_dims[3] = {0, 4, 4};
  CHECK(aloha_split_volume(bad_dims, 2, descs, 16, &cnt) != ALOHA_OK);
  CHECK(aloha_split_volume(dims, 5, descs, 2, &cnt) != ALOHA_OK);
  CHECK(aloha_volume_extract(raw.data(), dims, 3, &descs[0].voxel_range, vox[0].data(),
                             &descs[0].value_range) != ALOHA_OK);
  aloha_box3i outside{{0, 0, 0}, {dims[0] + 1, 2, 2}};
  CHECK(aloha_volume_extract(raw.data(), dims, 1, &outside, vox[0].data(), &descs[0].value_range) !=
        ALOHA_OK);
  // a truncated .vol file
  {
    FILE *f = std::fopen(vol.c_str(), "rb");
    std::vector<char> bytes(200);
    const size_t got = std::fread(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
    const std::string cut = dir + "/cut.vol";
    f = std::fopen(cut.c_str(), "wb");
    std::fwrite(bytes.data(), 1, got, f);
    std::fclose(f);
    CHECK(aloha_vol_read_header(cut.c_str(), &c2, &v2, &s2, back, 16, &n2) != ALOHA_OK);
  }
  // partitioning
  std::vector<aloha_box3f> doms(nb);
  for (int i = 0; i < nb; ++i) doms[i] = descs[i].space_range;
  for (int ranks : {1, 2, 3, nb}) {
    aloha_partitioner *p = nullptr;
    CHECK(aloha_partitioner_create_kd(doms.data(), nb, ranks, &p) == ALOHA_OK);
    int total = 0;
    for (int r = 0; r < ranks; ++r) {
      int32_t ids[16];
      CHECK(aloha_partitioner_rank_bricks(p, r, ids, 16, &cnt) == ALOHA_OK);
      total += cnt;
    }
    CHECK(total == nb);
    int32_t order[16];
    const float ref[3] = {-10.f, 50.f, 8.f};
    CHECK(aloha_partitioner_composite_order(p, ref, order, 16, &cnt) == ALOHA_OK && cnt == ranks);
    unsigned char nodes[16 * 64];
    CHECK(aloha_partitioner_kd_nodes(p, nodes, 64, &cnt) == ALOHA_OK);
    CHECK(aloha_partitioner_rank_bricks(p, ranks, nullptr, 0, &cnt) != ALOHA_OK);
    aloha_partitioner_destroy(p);
  }
  aloha_partitioner *rr = nullptr;
  CHECK(aloha_partitioner_create_rr(7, 3, &rr) == ALOHA_OK);
  aloha_partitioner_destroy(rr);
  CHECK(aloha_partitioner_create_kd(doms.data(), nb, 0, &rr) != ALOHA_OK);
  CHECK(aloha_partitioner_create_kd(doms.data(), 0, 2, &rr) == ALOHA_OK);  // no bricks: empty ranks
  CHECK(aloha_partitioner_rank_bricks(rr, 1, nullptr, 0, &cnt) == ALOHA_OK && cnt == 0);
  aloha_partitioner_destroy(rr);
  return 0;
}

static int synth_and_camera() {
  const int32_t dims[3] = {20, 14, 9};
  std::vector<float> f((size_t)dims[0] * dims[1] * dims[2]);
  CHECK(aloha_synth_fbm_f32(dims, 3, 6.f, 5, f.data()) == ALOHA_OK);
  CHECK(aloha_synth_blobs_f32(dims, 4, 5, 1.f, 3.f, .3f, 1.f, f.data()) == ALOHA_OK);
  std::vector<float> tf(4 * 256);
  CHECK(aloha_synth_tf(256, 6, 2, tf.data()) == ALOHA_OK);
  CHECK(aloha_synth_tf(0, 6, 2, tf.data()) != ALOHA_OK);
  aloha_box3f b{{0, 0, 0}, {20, 14, 9}};
  float cam[12];
  CHECK(aloha_synth_camera(&b, 64, 48, 45.f, 2.5f, 9, cam) == ALOHA_OK);
  int32_t rect[4];
  CHECK(aloha_project_boxes(cam, &b, 1, 64, 48, rect) == ALOHA_OK);
  CHECK(rect[0] <= rect[2] && rect[1] <= rect[3] && rect[2] <= 64 && rect[3] <= 48);
  aloha_box3f behind{{-1e6f, -1, -1}, {1e6f, 1, 1}};  // straddles the camera plane: whole image
  CHECK(aloha_project_boxes(cam, &behind, 1, 64, 48, rect) == ALOHA_OK);
  aloha_box3f nan_box{{NAN, 0, 0}, {1, 1, 1}};
  CHECK(aloha_project_boxes(cam, &nan_box, 1, 64, 48, rect) == ALOHA_OK);
  CHECK(rect[0] == 0 && rect[1] == 0 && rect[2] == 64 && rect[3] == 48);
  CHECK(aloha_project_boxes(cam, &b, 1, 0, 48, rect) != ALOHA_OK);
  return 0;
}

static int meshes(const std::string &dir) {
  aloha_box3f dom{{0, 0, 0}, {1, 1, 1}};
  aloha_mesh *m = nullptr;
  CHECK(aloha_mesh_synth_spheres(40, 6, 9, &dom, .05f, .15f, .1f, 7, &m) == ALOHA_OK);
  int32_t nc = 0;
  CHECK(aloha_mesh_split(m, 4, &nc) == ALOHA_OK && nc == 4);
  int64_t nv = 0, nt = 0;
  aloha_box3f bounds;
  CHECK(aloha_mesh_info(m, &nv, &nt, &bounds, &nc) == ALOHA_OK);
  std::vector<float> verts(3 * (size_t)nv);
  std::vector<int32_t> tris(3 * (size_t)nt);
  CHECK(aloha_mesh_get(m, verts.data(), tris.data()) == ALOHA_OK);
  int64_t ranges[2 * 8];
  aloha_box3f cb[8];
  CHECK(aloha_mesh_clusters(m, ranges, cb, 8, &nc) == ALOHA_OK && nc == 4);
  const std::string tri = dir + "/check.tri";
  CHECK(aloha_tri_write(tri.c_str(), m) == ALOHA_OK);
  aloha_mesh *m2 = nullptr;
  CHECK(aloha_tri_read(tri.c_str(), &m2) == ALOHA_OK);
  int64_t nv2 = 0, nt2 = 0;
  int32_t nc2 = 0;
  CHECK(aloha_mesh_info(m2, &nv2, &nt2, &bounds, &nc2) == ALOHA_OK && nt2 == nt && nc2 == nc);
  aloha_mesh_destroy(m2);
  // BVH of clusters 1 and 3 (one rank), sized by a first call
  const int32_t ids[2] = {1, 3};
  const uint32_t cols[2] = {5, 9};
  int64_t ntr = 0, nno = 0;
  int32_t depth = 0, root = 0;
  CHECK(aloha_tris_build_host(m, 2, ids, cols, &ntr, &nno, &depth, nullptr, nullptr, &root) ==
        ALOHA_OK);
  std::vector<unsigned char> nodes(128 * (size_t)nno);
  std::vector<float> dtris(12 * (size_t)ntr);
  CHECK(aloha_tris_build_host(m, 2, ids, cols, &ntr, &nno, &depth, nodes.data(), dtris.data(),
                              &root) == ALOHA_OK);
  CHECK(depth <= 32 && ntr > 0);
  const int32_t bad = 7;
  CHECK(aloha_tris_build_host(m, 1, &bad, cols, &ntr, &nno, &depth, nullptr, nullptr, &root) !=
        ALOHA_OK);
  // the rank's BVH and a camera for the oracle's walk
  float cam[12];
  CHECK(aloha_synth_camera(&dom, 48, 40, 40.f, 1.6f, 4, cam) == ALOHA_OK);
  FILE *f = std::fopen((dir + "/rank.bvh").c_str(), "wb");
  CHECK(f != nullptr);
  std::fwrite(&ntr, 8, 1, f);
  std::fwrite(&nno, 8, 1, f);
  std::fwrite(&root, 4, 1, f);
  std::fwrite(cam, 4, 12, f);
  std::fwrite(nodes.data(), 1, nodes.size(), f);
  std::fwrite(dtris.data(), 4, dtris.size(), f);
  std::fclose(f);
  aloha_mesh_destroy(m);
  // a truncated .tri file
  {
    FILE *in = std::fopen(tri.c_str(), "rb");
    std::vector<char> bytes(300);
    const size_t got = std::fread(bytes.data(), 1, bytes.size(), in);
    std::fclose(in);
    const std::string cut = dir + "/cut.tri";
    FILE *out = std::fopen(cut.c_str(), "wb");
    std::fwrite(bytes.data(), 1, got, out);
    std::fclose(out);
    aloha_mesh *m3 = nullptr;
    CHECK(aloha_tri_read(cut.c_str(), &m3) != ALOHA_OK);
  }
  // Wavefront OBJ: a quad (two triangles) and a triangle with negative indices
  const std::string obj = dir + "/check.obj";
  f = std::fopen(obj.c_str(), "w");
  std::fputs("# check\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
             "f 1/1/1 2/1/1 3/1/1 4/1/1\nv 0 0 1\nf -1 -4 -3\n", f);
  std::fclose(f);
  aloha_mesh *mo = nullptr;
  CHECK(aloha_mesh_read_obj(obj.c_str(), &mo) == ALOHA_OK);
  CHECK(aloha_mesh_info(mo, &nv, &nt, &bounds, &nc) == ALOHA_OK && nv == 5 && nt == 3);
  aloha_mesh_destroy(mo);
  CHECK(aloha_mesh_read_obj((dir + "/missing.obj").c_str(), &mo) != ALOHA_OK);
  const float v3[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
  const int32_t t_bad[3] = {0, 1, 3};
  CHECK(aloha_mesh_create(v3, 3, t_bad, 1, &mo) != ALOHA_OK);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  if (int rc = bricking(dir)) return rc;
  if (int rc = synth_and_camera()) return rc;
  if (int rc = meshes(dir)) return rc;
  std::printf("host ok\n");
  return 0;
}
