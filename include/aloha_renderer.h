/*
 * aloha_renderer.h -- the reference's application-level C++ API, kept so the
 * render-and-composite path drops in behind existing callers:
 *
 *   maui::mpi::DistributedRenderer   DistributedRenderer.h:29-71
 *   maui::Renderer                   waikikiRenderer.h:28-118
 *
 * Same class names, constructor and method signatures, event semantics
 * (setCamera/resize/screenShot/setColorMap only queue; processEvents() runs
 * the handlers in queue order) and statics (spp, heatMapEnabled,
 * heatMapScale).  Differences, all on one GPU:
 *   - commSize is the number of virtual ranks; one Renderer object holds all
 *     of them and its render() produces the composited frame;
 *   - processEvents() drains a process-local queue (no MPI);
 *   - fbPointer is device memory (as Headless allocates it in GPU mode);
 *   - dt defaults to the reference's hard-coded 2.0 and has a setter.
 * The owl vector types are replaced by layout-identical PODs.
 */
#pragma once

#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "aloha_amd.h"

namespace maui {

struct vec2i {
  int x, y;
  bool operator==(const vec2i &o) const { return x == o.x && y == o.y; }
  bool operator!=(const vec2i &o) const { return !(*this == o); }
};
struct vec3f {
  float x, y, z;
};
struct vec4f {
  float x, y, z, w;
};
struct box3f {
  vec3f lower, upper;
};

namespace mpi {

struct DistributedRenderer {
  DistributedRenderer();
  virtual ~DistributedRenderer();

  void setCamera(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du,
                 const vec3f &dir_dv);
  void resize(const vec2i &newSize);
  void screenShot(const std::string &baseName);
  void setColorMap(const std::vector<vec4f> &newCM);

  virtual void handleCameraEvent(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du,
                                 const vec3f &dir_dv) = 0;
  virtual void handleResizeEvent(const vec2i &newSize) = 0;
  virtual void handleScreenShotEvent(const std::string &baseName) {}
  virtual void handleColorMapEvent(const std::vector<vec4f> &newCM) {}

  // Replay queued events in order (DistributedRenderer.cpp:139-211).
  void processEvents();

  struct {
    int rank = 0;
    int size = 0;
  } commWorld;

  int gpuID = 0;

 private:
  struct Event;
  std::vector<std::unique_ptr<Event>> events_;
};

}  // namespace mpi

struct Renderer : mpi::DistributedRenderer {
  enum class Mode { TriMesh, Volume };

  // Loads a .vol brick file (Mode::Volume: bricks to commSize virtual ranks
  // with partitionKD) or a .tri cluster file (Mode::TriMesh: clusters in
  // blocks per rank, partitionRoundRobin) and uploads it
  // (waikikiRenderer.cpp:344-846).
  Renderer(int commRank, int commSize, int numIslands, const std::string &inFileName);
  // Same from bricks already in memory (x-fastest voxels, encoding dtype).
  Renderer(int commRank, int commSize, int numIslands, const box3f &modelBounds,
           const std::vector<aloha_brick_desc> &bricks, const std::vector<const void *> &voxels,
           int dtype);
  // Mode::TriMesh from a mesh in memory (split into clusters or not); the
  // mesh is not kept after the upload.
  Renderer(int commRank, int commSize, int numIslands, const aloha_mesh *mesh);
  ~Renderer();

  // One frame: every virtual rank marches its bricks (or traces its
  // clusters), the rank images are 'over'-composited in visibility order
  // (depth-composited for meshes) and packed to RGBA8 into the device buffer
  // fbPointer (waikikiRenderer.cpp:1015-1080).
  void render(uint32_t *fbPointer);

  void handleCameraEvent(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du,
                         const vec3f &dir_dv) override;
  void handleResizeEvent(const vec2i &newSize) override;
  void handleScreenShotEvent(const std::string &baseName) override;
  void handleColorMapEvent(const std::vector<vec4f> &newCM) override;

  void resetAccum() { accumID = 0; }
  void setDt(float dt) { this->dt = dt; }
  // the last composited frame as float RGBA (device, fbSize.x*fbSize.y*4;
  // Mode::Volume -- the depth composite of meshes is RGBA8 only)
  const float *compositeRGBA() const { return d_comp_; }
  // the last depth-composited depth buffer (device, Mode::TriMesh)
  const float *compositeDepth() const { return d_comp_depth_; }
  // status of the last failing call ("" if none)
  const std::string &lastError() const { return error_; }

  Mode mode = Mode::Volume;
  box3f modelBounds{};
  std::vector<int> compositeOrder;            // rank ids, front to back
  std::vector<std::vector<int>> perRank;      // brick (cluster) ids per rank
  int accumID = 0;
  float dt = 2.f;
  vec2i fbSize{1, 1};
  uint32_t *fbPointer = nullptr;
  float maxDepth = 1.f;
  int islandID = 0;
  int numIslands = 1;
  int displayRank = 0;
  // The reference's data path (the defaults): each rank's frame is packed to
  // RGBA8 with the crosshair inverted on row H/2 and column W/2
  // (deviceCode.cpp:410-416), and IceT blends those bytes (RGBA_UBYTE,
  // waikikiRenderer.cpp:908-916).  ubyteComposite = false blends the ranks'
  // float images and packs once (up to 2/255 apart from the reference's
  // bytes); crossHairs = false leaves the rank images without the crosshair.
  // compositeRGBA() is the float blend either way.
  bool ubyteComposite = true;
  bool crossHairs = true;

  static int spp;
  static bool heatMapEnabled;
  static float heatMapScale;

 private:
  void init(int commRank, int commSize, int numIslands,
            const std::vector<aloha_brick_desc> &bricks, const std::vector<const void *> &voxels,
            int dtype);
  void initRanks(int commRank, int commSize, int numIslands);
  void initMesh(int commRank, int commSize, int numIslands, const aloha_mesh *mesh);
  void check(int rc);
  void freeImages();
  aloha_ctx *ctx_ = nullptr;
  aloha_partitioner *part_ = nullptr;
  std::vector<int> brickIds_;
  std::vector<int32_t> geomIds_;  // Mode::TriMesh: one geometry per rank
  int islandSize_ = 1;
  float *d_rank_rgba_ = nullptr;
  uint32_t *d_rank_rgba8_ = nullptr;
  float *d_rank_depth_ = nullptr;
  float *d_rank_accum_ = nullptr;
  float *d_comp_ = nullptr;
  float *d_comp_depth_ = nullptr;
  std::string error_;
};

}  // namespace maui
