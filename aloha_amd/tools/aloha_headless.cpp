// aloha_headless -- renders frames through the maui::Renderer C++ API, like
// the reference's non-display ranks (Headless.cpp) but with every virtual
// rank on this GPU, and writes the last frame.
//
//   aloha_headless in.vol|in.tri -ranks P [-islands I] [-rank R] [-size W H]
//                  [-camera ox oy oz d00x d00y d00z dux duy duz dvx dvy dvz]
//                  [-tf tf.f32 (n x RGBA float; volumes)] [-dt dt] [-spp n]
//                  [-frames F] [-float-composite] [-no-crosshair] [-o out]
//     -> out.rgba8 (W*H u32) and out.rgba (W*H*4 f32) for volumes,
//        out.rgba8 and out.depth (W*H f32) for meshes
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "aloha_renderer.h"

int main(int argc, char **argv) {
  std::string in, out = "frame", tf_path;
  int P = 1, islands = 1, rank = 0, frames = 1, W = 512, H = 512;
  float dt = 2.f;
  bool ubyte = true, crosshair = true;  // the reference's frame (Renderer defaults)
  std::vector<float> cam;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a[0] != '-') in = a;
    else if (a == "-ranks") P = std::atoi(argv[++i]);
    else if (a == "-islands") islands = std::atoi(argv[++i]);
    else if (a == "-rank") rank = std::atoi(argv[++i]);
    else if (a == "-frames") frames = std::atoi(argv[++i]);
    else if (a == "-spp") maui::Renderer::spp = std::atoi(argv[++i]);
    else if (a == "-dt") dt = std::atof(argv[++i]);
    else if (a == "-tf") tf_path = argv[++i];
    else if (a == "-float-composite") ubyte = false;
    else if (a == "-no-crosshair") crosshair = false;
    else if (a == "-o") out = argv[++i];
    else if (a == "-size") {
      W = std::atoi(argv[++i]);
      H = std::atoi(argv[++i]);
    } else if (a == "-camera") {
      for (int k = 0; k < 12; ++k) cam.push_back(std::atof(argv[++i]));
    } else {
      fprintf(stderr, "unknown arg %s\n", a.c_str());
      return 1;
    }
  }
  const bool mesh = in.size() >= 4 && in.compare(in.size() - 4, 4, ".tri") == 0;
  if (in.empty() || cam.size() != 12 || (tf_path.empty() && !mesh)) {
    fprintf(stderr, "usage: aloha_headless in.vol|in.tri -ranks P -camera <12 floats> "
                    "-tf tf.f32 ...\n");
    return 1;
  }
  try {
    maui::Renderer r(rank, P, islands, in);
    r.setDt(dt);
    r.ubyteComposite = ubyte;
    r.crossHairs = crosshair;
    r.resize({W, H});
    if (!tf_path.empty()) {
      std::ifstream tf(tf_path, std::ios::binary | std::ios::ate);
      std::vector<maui::vec4f> cm((size_t)tf.tellg() / sizeof(maui::vec4f));
      tf.seekg(0);
      tf.read((char *)cm.data(), (std::streamsize)(cm.size() * sizeof(maui::vec4f)));
      r.setColorMap(cm);
    }
    r.setCamera({cam[0], cam[1], cam[2]}, {cam[3], cam[4], cam[5]}, {cam[6], cam[7], cam[8]},
                {cam[9], cam[10], cam[11]});
    uint32_t *fb = nullptr;
    if (hipMalloc(&fb, (size_t)W * H * 4) != hipSuccess) return 2;
    for (int f = 0; f < frames; ++f) r.render(fb);
    std::vector<uint32_t> px((size_t)W * H);
    (void)hipMemcpy(px.data(), fb, px.size() * 4, hipMemcpyDeviceToHost);
    std::ofstream(out + ".rgba8", std::ios::binary).write((const char *)px.data(), (std::streamsize)(px.size() * 4));
    if (mesh) {
      std::vector<float> depth((size_t)W * H);
      (void)hipMemcpy(depth.data(), r.compositeDepth(), depth.size() * 4, hipMemcpyDeviceToHost);
      std::ofstream(out + ".depth", std::ios::binary).write((const char *)depth.data(), (std::streamsize)(depth.size() * 4));
    } else {
      std::vector<float> rgba((size_t)W * H * 4);
      (void)hipMemcpy(rgba.data(), r.compositeRGBA(), rgba.size() * 4, hipMemcpyDeviceToHost);
      std::ofstream(out + ".rgba", std::ios::binary).write((const char *)rgba.data(), (std::streamsize)(rgba.size() * 4));
    }
    printf("rendered %d frame(s) %dx%d, %d ranks, order", frames, W, H, P);
    for (int k : r.compositeOrder) printf(" %d", k);
    printf("\n");
    (void)hipFree(fb);
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
  return 0;
}
