// aloha_chopsuey -- chopSuey (chopSuey.cpp:361-440) over the aloha_amd C ABI:
// bricks a raw volume into a .vol file (VolumeSplitter, :232-359), or splits
// an .obj mesh into clusters and writes a .tri file (MeshSplitter, :48-229).
//
//   aloha_chopsuey in.raw -dims X Y Z [-type uint8|uint16|float | -bpc N]
//                  [-n numBricks] [-o out.vol]
//   aloha_chopsuey in.obj [-n numClusters] [-o out.tri]
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "aloha_amd.h"

static int usage(const std::string &err) {
  if (!err.empty()) fprintf(stderr, "\nFatal error: %s\n\n", err.c_str());
  fprintf(stderr, "Usage: ./aloha_chopsuey inFile.raw -dims x y z -type uint8|uint16|float "
                  "-n numClusters -o outFile.vol\n");
  return 1;
}

int main(int argc, char **argv) {
  std::string in, out = "chopSuey.vol";
  int n = 1, bpc = 1;
  int32_t dims[3] = {0, 0, 0};
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int k) { return i + k < argc; };
    if (a[0] != '-') in = a;
    else if (a == "-o" && need(1)) out = argv[++i];
    else if (a == "-n" && need(1)) n = std::atoi(argv[++i]);
    else if (a == "-dims" && need(3)) {
      for (int k = 0; k < 3; ++k) dims[k] = std::atoi(argv[++i]);
    } else if (a == "-bpc" && need(1)) bpc = std::atoi(argv[++i]);
    else if (a == "-type" && need(1)) {
      const std::string t = argv[++i];
      if (t == "uint8") bpc = 1;
      else if (t == "uint16") bpc = 2;
      else if (t == "float") bpc = 4;
      else return usage("wrong type '" + t + "'");
    } else return usage("unknown cmdline arg '" + a + "'");
  }
  if (in.empty()) return usage("no filename specified");
  if (in.size() >= 4 && in.compare(in.size() - 4, 4, ".obj") == 0) {
    aloha_mesh *mesh = nullptr;
    int32_t count = 0;
    if (aloha_mesh_read_obj(in.c_str(), &mesh)) return usage(aloha_last_error(nullptr));
    if (aloha_mesh_split(mesh, n, &count) || aloha_tri_write(out.c_str(), mesh)) {
      aloha_mesh_destroy(mesh);
      return usage(aloha_last_error(nullptr));
    }
    std::vector<int64_t> ranges(2 * (size_t)count);
    std::vector<aloha_box3f> bounds(count);
    aloha_mesh_clusters(mesh, ranges.data(), bounds.data(), count, &count);
    for (int i = 0; i < count; ++i)
      printf("%lld %lld [(%g,%g,%g),(%g,%g,%g)]\n", (long long)ranges[2 * i],
             (long long)ranges[2 * i + 1], bounds[i].lower[0], bounds[i].lower[1],
             bounds[i].lower[2], bounds[i].upper[0], bounds[i].upper[1], bounds[i].upper[2]);
    aloha_mesh_destroy(mesh);
    return 0;
  }
  if (!dims[0] || !dims[1] || !dims[2]) return usage("no input dimensions specified");
  if (bpc != 1 && bpc != 2 && bpc != 4) return usage("no data type/bpc specified");
  const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<char> raw(nvox * bpc);
  std::ifstream f(in, std::ios::binary);
  if (!f.read(raw.data(), (std::streamsize)raw.size())) return usage("cannot read " + in);

  std::vector<aloha_brick_desc> bricks(4 * (size_t)n + 8);
  int32_t count = 0;
  if (aloha_split_volume(dims, n, bricks.data(), (int32_t)bricks.size(), &count))
    return usage(aloha_last_error(nullptr));
  bricks.resize(count);
  std::vector<std::vector<float>> vox(count);
  std::vector<const float *> ptrs(count);
  for (int i = 0; i < count; ++i) {
    size_t nv = 1;
    for (int a = 0; a < 3; ++a)
      nv *= (size_t)(bricks[i].voxel_range.upper[a] - bricks[i].voxel_range.lower[a]);
    vox[i].resize(nv);
    bricks[i].value_range = {FLT_MAX, -FLT_MAX};
    if (aloha_volume_extract(raw.data(), dims, bpc, &bricks[i].voxel_range, vox[i].data(),
                             &bricks[i].value_range))
      return usage(aloha_last_error(nullptr));
    ptrs[i] = vox[i].data();
    printf("%d [%g,%g]\n", i, bricks[i].value_range.lower, bricks[i].value_range.upper);
  }
  aloha_box3i cell{{0, 0, 0}, {dims[0], dims[1], dims[2]}};
  aloha_box3i voxr{{0, 0, 0}, {dims[0] + 1, dims[1] + 1, dims[2] + 1}};
  aloha_box3f space{{0, 0, 0}, {(float)dims[0], (float)dims[1], (float)dims[2]}};
  if (aloha_vol_write(out.c_str(), &cell, &voxr, &space, count, bricks.data(), ptrs.data()))
    return usage(aloha_last_error(nullptr));
  return 0;
}
