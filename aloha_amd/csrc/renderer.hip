// renderer.hip -- maui::Renderer / maui::mpi::DistributedRenderer on top of
// the C ABI (waikikiRenderer.cpp, DistributedRenderer.cpp).  All virtual
// ranks live in this process and on this GPU: the MPI event broadcast
// becomes a local queue, the per-rank OWL launches become one
// aloha_render_ranks launch per island, and IceT becomes aloha_composite_over_sparse.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "aloha_renderer.h"

namespace aloha {
// the library's allocation entry (context.hip): hipMalloc behind the
// allocation-fault test hook
hipError_t dev_malloc(void **p, size_t bytes);
}  // namespace aloha

namespace maui {

int Renderer::spp = 1;
bool Renderer::heatMapEnabled = false;
float Renderer::heatMapScale = 1e-5f;

// ---------------------------------------------------------------------------
// DistributedRenderer: event queue (DistributedRenderer.cpp:93-211)
// ---------------------------------------------------------------------------
namespace mpi {

struct DistributedRenderer::Event {
  enum class Type { Camera, Resize, ScreenShot, ColorMap } type;
  vec3f cam[4];
  vec2i size;
  std::string name;
  std::vector<vec4f> cm;
};

DistributedRenderer::DistributedRenderer() {
  commWorld.rank = 0;
  commWorld.size = 1;
}

DistributedRenderer::~DistributedRenderer() = default;

void DistributedRenderer::setCamera(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du,
                                    const vec3f &dir_dv) {
  auto ev = std::make_unique<Event>();
  ev->type = Event::Type::Camera;
  ev->cam[0] = org, ev->cam[1] = dir_00, ev->cam[2] = dir_du, ev->cam[3] = dir_dv;
  events_.push_back(std::move(ev));
}

void DistributedRenderer::resize(const vec2i &newSize) {
  auto ev = std::make_unique<Event>();
  ev->type = Event::Type::Resize;
  ev->size = newSize;
  events_.push_back(std::move(ev));
}

void DistributedRenderer::screenShot(const std::string &baseName) {
  auto ev = std::make_unique<Event>();
  ev->type = Event::Type::ScreenShot;
  ev->name = baseName;
  events_.push_back(std::move(ev));
}

void DistributedRenderer::setColorMap(const std::vector<vec4f> &newCM) {
  auto ev = std::make_unique<Event>();
  ev->type = Event::Type::ColorMap;
  ev->cm = newCM;
  events_.push_back(std::move(ev));
}

void DistributedRenderer::processEvents() {
  // handlers may queue new events; those run on the next call, as in the
  // reference where they are broadcast on the next frame
  std::vector<std::unique_ptr<Event>> evs;
  evs.swap(events_);
  for (auto &ev : evs) {
    switch (ev->type) {
      case Event::Type::Camera:
        handleCameraEvent(ev->cam[0], ev->cam[1], ev->cam[2], ev->cam[3]);
        break;
      case Event::Type::Resize:
        handleResizeEvent(ev->size);
        break;
      case Event::Type::ScreenShot:
        handleScreenShotEvent(ev->name);
        break;
      case Event::Type::ColorMap:
        handleColorMapEvent(ev->cm);
        break;
    }
  }
}

}  // namespace mpi

// ---------------------------------------------------------------------------
// Renderer
// ---------------------------------------------------------------------------
void Renderer::check(int rc) {
  if (rc != ALOHA_OK) {
    error_ = aloha_last_error(ctx_);
    throw std::runtime_error("aloha: " + error_);
  }
}

Renderer::Renderer(int commRank, int commSize, int numIslands, const std::string &inFileName) {
  // mode from the file extension (waikikiRenderer.cpp:390-394)
  if (inFileName.size() >= 4 && inFileName.compare(inFileName.size() - 4, 4, ".tri") == 0) {
    aloha_mesh *mesh = nullptr;
    check(aloha_tri_read(inFileName.c_str(), &mesh));
    try {
      initMesh(commRank, commSize, numIslands, mesh);
    } catch (...) {
      aloha_mesh_destroy(mesh);
      throw;
    }
    aloha_mesh_destroy(mesh);
    return;
  }
  aloha_box3i cr, vr;
  aloha_box3f sr;
  int32_t n = 0;
  check(aloha_vol_read_header(inFileName.c_str(), &cr, &vr, &sr, nullptr, 0, &n));
  std::vector<aloha_brick_desc> bricks(n);
  check(aloha_vol_read_header(inFileName.c_str(), &cr, &vr, &sr, bricks.data(), n, &n));
  std::vector<std::vector<float>> storage(n);
  std::vector<const void *> voxels(n);
  for (int i = 0; i < n; ++i) {
    size_t nv = 1;
    for (int a = 0; a < 3; ++a)
      nv *= (size_t)(bricks[i].voxel_range.upper[a] - bricks[i].voxel_range.lower[a]);
    storage[i].resize(nv);
    check(aloha_vol_read_brick(inFileName.c_str(), i, storage[i].data(), nv));
    voxels[i] = storage[i].data();
  }
  modelBounds = {{sr.lower[0], sr.lower[1], sr.lower[2]}, {sr.upper[0], sr.upper[1], sr.upper[2]}};
  init(commRank, commSize, numIslands, bricks, voxels, ALOHA_F32);
}

Renderer::Renderer(int commRank, int commSize, int numIslands, const box3f &bounds,
                   const std::vector<aloha_brick_desc> &bricks,
                   const std::vector<const void *> &voxels, int dtype) {
  modelBounds = bounds;
  init(commRank, commSize, numIslands, bricks, voxels, dtype);
}

Renderer::Renderer(int commRank, int commSize, int numIslands, const aloha_mesh *mesh) {
  initMesh(commRank, commSize, numIslands, mesh);
}

void Renderer::initRanks(int commRank, int commSize, int nIslands) {
  if (nIslands < 1 || commSize < nIslands || commSize % nIslands != 0)
    throw std::invalid_argument("aloha: commSize must be a multiple of numIslands");
  commWorld.rank = commRank;
  commWorld.size = commSize;
  numIslands = nIslands;
  islandSize_ = commSize / nIslands;
  islandID = commRank / islandSize_;
  check(aloha_ctx_create(0, &ctx_));
}

// TriMesh set-up (waikikiRenderer.cpp:396-593): clusters in blocks per rank
// (partitionRoundRobin), rank r's i-th cluster coloured by
// randomColor(r * (numClusters / ranks) + i), one BVH per rank.
void Renderer::initMesh(int commRank, int commSize, int nIslands, const aloha_mesh *mesh) {
  if (!mesh) throw std::invalid_argument("aloha: mesh is NULL");
  initRanks(commRank, commSize, nIslands);
  mode = Mode::TriMesh;
  aloha_box3f b;
  int32_t nc = 0;
  check(aloha_mesh_info(mesh, nullptr, nullptr, &b, &nc));
  modelBounds = {{b.lower[0], b.lower[1], b.lower[2]}, {b.upper[0], b.upper[1], b.upper[2]}};
  const int n = nc > 0 ? nc : 1;
  check(aloha_partitioner_create_rr(n, islandSize_, &part_));
  perRank.assign(islandSize_, {});
  std::vector<int32_t> offs{0}, ids;
  std::vector<uint32_t> colors;
  const int per = n / islandSize_;
  for (int r = 0; r < islandSize_; ++r) {
    int32_t cnt = 0;
    std::vector<int32_t> got(n + 1);
    check(aloha_partitioner_rank_bricks(part_, r, got.data(), (int32_t)got.size(), &cnt));
    perRank[r].assign(got.begin(), got.begin() + cnt);
    for (int i = 0; i < cnt; ++i) {
      ids.push_back(got[i]);
      colors.push_back((uint32_t)(r * per + i));
    }
    offs.push_back((int32_t)ids.size());
  }
  geomIds_.resize(islandSize_);
  check(aloha_tris_upload(ctx_, mesh, islandSize_, offs.data(), ids.data(), colors.data(),
                          geomIds_.data()));
  compositeOrder.resize(islandSize_);
  for (int r = 0; r < islandSize_; ++r) compositeOrder[r] = r;
}

void Renderer::init(int commRank, int commSize, int nIslands,
                    const std::vector<aloha_brick_desc> &bricks,
                    const std::vector<const void *> &voxels, int dtype) {
  initRanks(commRank, commSize, nIslands);
  mode = Mode::Volume;
  std::vector<aloha_box3f> domains(bricks.size());
  for (size_t i = 0; i < bricks.size(); ++i) domains[i] = bricks[i].space_range;
  check(aloha_partitioner_create_kd(domains.data(), (int32_t)domains.size(), islandSize_, &part_));
  perRank.assign(islandSize_, {});
  for (int r = 0; r < islandSize_; ++r) {
    int32_t cnt = 0;
    std::vector<int32_t> ids(bricks.size() + 1);
    check(aloha_partitioner_rank_bricks(part_, r, ids.data(), (int32_t)ids.size(), &cnt));
    perRank[r].assign(ids.begin(), ids.begin() + cnt);
  }
  brickIds_.resize(bricks.size());
  for (size_t i = 0; i < bricks.size(); ++i)
    check(aloha_brick_upload(ctx_, &bricks[i], voxels[i], dtype, 0, &brickIds_[i]));
  compositeOrder.resize(islandSize_);
  for (int r = 0; r < islandSize_; ++r) compositeOrder[r] = r;
}

Renderer::~Renderer() {
  freeImages();
  if (part_) aloha_partitioner_destroy(part_);
  if (ctx_) aloha_ctx_destroy(ctx_);
}

void Renderer::freeImages() {
  for (void *p : {(void *)d_rank_rgba_, (void *)d_rank_rgba8_, (void *)d_rank_depth_,
                  (void *)d_rank_accum_, (void *)d_comp_, (void *)d_comp_depth_})
    if (p) (void)hipFree(p);
  d_rank_rgba_ = d_rank_accum_ = d_comp_ = d_rank_depth_ = d_comp_depth_ = nullptr;
  d_rank_rgba8_ = nullptr;
}

void Renderer::handleCameraEvent(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du,
                                 const vec3f &dir_dv) {
  // maxDepth = length(org + center) + length(center + upper) (:858-860)
  const vec3f c = {(modelBounds.lower.x + modelBounds.upper.x) * .5f,
                   (modelBounds.lower.y + modelBounds.upper.y) * .5f,
                   (modelBounds.lower.z + modelBounds.upper.z) * .5f};
  auto length = [](float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); };
  const float radius =
      length(c.x + modelBounds.upper.x, c.y + modelBounds.upper.y, c.z + modelBounds.upper.z);
  maxDepth = length(org.x + c.x, org.y + c.y, org.z + c.z) + radius;
  const float o[3] = {org.x, org.y, org.z}, d00[3] = {dir_00.x, dir_00.y, dir_00.z},
              du[3] = {dir_du.x, dir_du.y, dir_du.z}, dv[3] = {dir_dv.x, dir_dv.y, dir_dv.z};
  check(aloha_camera_set(ctx_, o, d00, du, dv));
  std::vector<int32_t> order(islandSize_ + 1);
  int32_t cnt = 0;
  check(aloha_partitioner_composite_order(part_, o, order.data(), (int32_t)order.size(), &cnt));
  compositeOrder.assign(order.begin(), order.begin() + cnt);
  accumID = 0;
}

void Renderer::handleResizeEvent(const vec2i &newSize) {
  if (newSize.x < 1 || newSize.y < 1) throw std::invalid_argument("aloha: empty frame size");
  if (newSize == fbSize && d_comp_) return;
  freeImages();
  fbSize = newSize;
  const size_t npix = (size_t)newSize.x * newSize.y;
  const size_t P = (size_t)commWorld.size;
  auto alloc = [&](void **p, size_t bytes) {
    hipError_t e = aloha::dev_malloc(p, bytes);
    if (e != hipSuccess) throw std::runtime_error(std::string("aloha: hipMalloc: ") + hipGetErrorString(e));
  };
  alloc((void **)&d_rank_rgba_, P * npix * 16);
  alloc((void **)&d_rank_rgba8_, P * npix * 4);
  alloc((void **)&d_rank_depth_, P * npix * 4);
  alloc((void **)&d_rank_accum_, P * npix * 16);
  alloc((void **)&d_comp_, npix * 16);
  alloc((void **)&d_comp_depth_, npix * 4);
  accumID = 0;
}

void Renderer::handleColorMapEvent(const std::vector<vec4f> &newCM) {
  check(aloha_tf_set(ctx_, (const float *)newCM.data(), (int32_t)newCM.size()));
  accumID = 0;
}

namespace {
// Minimal PNG writer (stored deflate blocks, no compression).
uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xffffffffu) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t r = i;
      for (int k = 0; k < 8; ++k) r = (r & 1) ? 0xedb88320u ^ (r >> 1) : r >> 1;
      table[i] = r;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 255] ^ (c >> 8);
  return c;
}
void put32(std::vector<uint8_t> &v, uint32_t x) {
  for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}
void chunk(std::ofstream &f, const char *type, const std::vector<uint8_t> &data) {
  std::vector<uint8_t> b;
  put32(b, (uint32_t)data.size());
  b.insert(b.end(), type, type + 4);
  b.insert(b.end(), data.begin(), data.end());
  put32(b, crc32(b.data() + 4, b.size() - 4) ^ 0xffffffffu);
  f.write((const char *)b.data(), (std::streamsize)b.size());
}
bool write_png(const std::string &path, int W, int H, const std::vector<uint32_t> &rgba) {
  std::ofstream f(path, std::ios::binary);
  if (!f) return false;
  const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  f.write((const char *)sig, 8);
  std::vector<uint8_t> ihdr;
  put32(ihdr, (uint32_t)W);
  put32(ihdr, (uint32_t)H);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
  chunk(f, "IHDR", ihdr);
  std::vector<uint8_t> raw;
  for (int y = 0; y < H; ++y) {
    raw.push_back(0);
    // row 0 of the frame buffer is the bottom row (viewer lower_left origin)
    const uint32_t *line = rgba.data() + (size_t)(H - 1 - y) * W;
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 4; ++c) raw.push_back(c == 3 ? 0xff : (uint8_t)(line[x] >> (8 * c)));
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t v : raw) a = (a + v) % 65521, b = (b + a) % 65521;
  for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
    const size_t len = std::min<size_t>(65535, raw.size() - off);
    z.push_back(off + len >= raw.size() ? 1 : 0);
    z.push_back((uint8_t)len), z.push_back((uint8_t)(len >> 8));
    z.push_back((uint8_t)~len), z.push_back((uint8_t)(~len >> 8));
    z.insert(z.end(), raw.begin() + (long)off, raw.begin() + (long)(off + len));
    if (len == 0) break;
  }
  put32(z, (b << 16) | a);
  chunk(f, "IDAT", z);
  chunk(f, "IEND", {});
  return (bool)f;
}
}  // namespace

void Renderer::handleScreenShotEvent(const std::string &baseName) {
  if (!fbPointer) return;
  std::vector<uint32_t> px((size_t)fbSize.x * fbSize.y);
  if (hipMemcpy(px.data(), fbPointer, px.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return;
  const std::string name = baseName + "_rank" + std::to_string(commWorld.rank % islandSize_) +
                           "_island" + std::to_string(islandID) + ".png";
  write_png(name, fbSize.x, fbSize.y, px);
}

void Renderer::render(uint32_t *fb) {
  fbPointer = fb;
  processEvents();
  if (!d_comp_) throw std::logic_error("aloha: render before resize");
  aloha_render_params p{};
  p.dt = dt;
  p.max_depth = maxDepth;
  p.spp = spp > 1 ? spp : 1;
  p.accum_id = accumID++;
  p.num_islands = numIslands;
  p.jitter_mode = ALOHA_JITTER_RANDOM;
  p.flags = (heatMapEnabled ? ALOHA_FLAG_HEATMAP : 0) | (crossHairs ? ALOHA_FLAG_CROSSHAIR : 0);
  p.heat_map_scale = heatMapScale;
  const int W = fbSize.x, H = fbSize.y;
  const size_t npix = (size_t)W * H;
  std::vector<int32_t> offs{0}, ids;
  if (mode == Mode::Volume)
    for (int r = 0; r < islandSize_; ++r) {
      for (int b : perRank[r]) ids.push_back(brickIds_[b]);
      offs.push_back((int32_t)ids.size());
    }
  for (int i = 0; i < numIslands; ++i) {
    p.island_id = i;
    const size_t base = (size_t)i * islandSize_ * npix;
    aloha_images img{d_rank_rgba_ + 4 * base, d_rank_rgba8_ + base, d_rank_depth_ + base,
                     d_rank_accum_ + 4 * base, nullptr};
    if (mode == Mode::TriMesh)
      check(aloha_render_tris(ctx_, &p, W, H, islandSize_, geomIds_.data(), &img, nullptr));
    else
      check(aloha_render_ranks(ctx_, &p, W, H, islandSize_, offs.data(), ids.data(), &img));
  }
  if (mode == Mode::TriMesh) {
    // IceT Z_BUFFER / REDUCE with an RGBA8 colour and float depth (:903-909)
    check(aloha_composite_zmin(ctx_, d_rank_rgba8_, d_rank_depth_, commWorld.size, W, H,
                               commWorld.rank == displayRank ? fb : nullptr, d_comp_depth_));
  } else {
    // composite order expanded over islands (:1043-1050); each rank image is
    // read only inside the rectangle its bricks project to (transparent
    // outside it; the whole image while frames accumulate)
    std::vector<int32_t> order;
    for (int i = 0; i < numIslands; ++i)
      for (int r : compositeOrder) order.push_back(i * islandSize_ + r);
    std::vector<int32_t> rects(4 * (size_t)islandSize_);
    check(aloha_rank_rects(ctx_, &p, W, H, islandSize_, offs.data(), ids.data(), rects.data()));
    std::vector<int32_t> all;
    for (int i = 0; i < numIslands; ++i) all.insert(all.end(), rects.begin(), rects.end());
    const bool display = commWorld.rank == displayRank;
    check(aloha_composite_over_sparse(ctx_, d_rank_rgba_, commWorld.size, order.data(),
                                      all.data(), W, H, d_comp_,
                                      display && !ubyteComposite ? fb : nullptr));
    // IceT's RGBA_UBYTE blend of the ranks' 8-bit frames (:908-916, :1066-1077)
    if (display && ubyteComposite)
      check(aloha_composite_over_rgba8(ctx_, d_rank_rgba8_, commWorld.size, order.data(), W, H,
                                       fb));
  }
  check(aloha_sync(ctx_));
  if (commWorld.rank != displayRank) {
    // IceT hands non-display ranks no image; their fb keeps their own frame
    const hipError_t e = hipMemcpy(fb, d_rank_rgba8_ + (size_t)commWorld.rank * npix, npix * 4,
                                   hipMemcpyDeviceToDevice);
    if (e != hipSuccess) throw std::runtime_error(std::string("aloha: ") + hipGetErrorString(e));
  }
}

}  // namespace maui
