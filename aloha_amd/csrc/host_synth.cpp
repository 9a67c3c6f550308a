// host_synth.cpp -- seeded synthetic inputs of the benchmark workloads:
// Gaussian-blob and fBm volumes, a piecewise-linear transfer function and an
// orbit camera in the viewer's (org, dir_00, dir_du, dir_dv) convention
// (viewer.cpp:148-158).  Random numbers come from splitmix64 so the inputs
// do not depend on the C++ library's distribution implementations.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "internal.h"

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
  // uniform double in [0,1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

constexpr unsigned kMaxThreads = 64;

// Run f(thread, z0, z1) over z slabs on up to kMaxThreads hardware threads.
template <typename F>
void parallel_z(int nz, F f) {
  unsigned nt = std::max(1u, std::min(kMaxThreads, std::thread::hardware_concurrency()));
  nt = std::min<unsigned>(nt, (unsigned)std::max(1, nz));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    const int z0 = (int)((long long)nz * t / nt), z1 = (int)((long long)nz * (t + 1) / nt);
    th.emplace_back([=] { f(t, z0, z1); });
  }
  for (auto &x : th) x.join();
}

struct Blob {
  double c[3], inv2s2, amp;
};

std::vector<Blob> make_blobs(const int32_t dims[3], int n, uint64_t seed, float smin, float smax,
                             float amin, float amax) {
  SplitMix64 rng(seed);
  std::vector<Blob> b(n);
  for (auto &x : b) {
    for (int a = 0; a < 3; ++a) x.c[a] = rng.uniform(0.0, (double)dims[a]);
    const double s = rng.uniform(smin, smax);
    x.inv2s2 = 1.0 / (2.0 * s * s);
    x.amp = rng.uniform(amin, amax);
  }
  return b;
}

// Sum of blobs at voxel centres (integer coordinates), double precision.
std::vector<double> blob_field(const int32_t dims[3], const std::vector<Blob> &blobs) {
  const size_t nx = dims[0], ny = dims[1], nz = dims[2];
  std::vector<double> f(nx * ny * nz);
  parallel_z((int)nz, [&](unsigned, int z0, int z1) {
    for (int z = z0; z < z1; ++z)
      for (size_t y = 0; y < ny; ++y) {
        double *row = &f[((size_t)z * ny + y) * nx];
        for (const Blob &b : blobs) {
          const double dy = (double)y - b.c[1], dz = (double)z - b.c[2];
          const double eyz = b.amp * std::exp(-(dy * dy + dz * dz) * b.inv2s2);
          if (eyz < 1e-300) continue;
          for (size_t x = 0; x < nx; ++x) {
            const double dx = (double)x - b.c[0];
            row[x] += eyz * std::exp(-dx * dx * b.inv2s2);
          }
        }
      }
  });
  return f;
}

// --- gradient noise -------------------------------------------------------
inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  uint32_t h = seed ^ 0x27d4eb2fu;
  h ^= (uint32_t)x * 0x8da6b343u;
  h ^= (uint32_t)y * 0xd8163841u;
  h ^= (uint32_t)z * 0xcb1ab31fu;
  h ^= h >> 15;
  h *= 0x2c1b3c6du;
  h ^= h >> 12;
  h *= 0x297a2d39u;
  h ^= h >> 15;
  return h;
}

// one of the 12 cube-edge gradients dotted with (x,y,z)
inline double grad_dot(uint32_t h, double x, double y, double z) {
  switch (h % 12u) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    default: return -y - z;
  }
}

inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

double gradient_noise(double px, double py, double pz, uint32_t seed) {
  const double fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
  const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  const double x = px - fx, y = py - fy, z = pz - fz;
  const double u = fade(x), v = fade(y), w = fade(z);
  double c[2][2][2];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i)
        c[k][j][i] = grad_dot(hash3(ix + i, iy + j, iz + k, seed), x - i, y - j, z - k);
  auto L = [](double a, double b, double t) { return a + t * (b - a); };
  const double x00 = L(c[0][0][0], c[0][0][1], u), x10 = L(c[0][1][0], c[0][1][1], u);
  const double x01 = L(c[1][0][0], c[1][0][1], u), x11 = L(c[1][1][0], c[1][1][1], u);
  return L(L(x00, x10, v), L(x01, x11, v), w);
}

}  // namespace

extern "C" int aloha_synth_blobs_u8(const int32_t dims[3], int32_t num_blobs, uint64_t seed,
                                    float smin, float smax, float amin, float amax,
                                    uint8_t *out) {
  if (!dims || !out || num_blobs < 1 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_synth_blobs_u8: bad argument");
  const auto f = blob_field(dims, make_blobs(dims, num_blobs, seed, smin, smax, amin, amax));
  double mx = 0.0;
  for (double v : f) mx = std::max(mx, v);
  if (!(mx > 0.0)) mx = 1.0;
  for (size_t i = 0; i < f.size(); ++i) {
    const double q = std::nearbyint(255.0 * f[i] / mx);
    out[i] = (uint8_t)std::min(255.0, std::max(0.0, q));
  }
  return ALOHA_OK;
}

extern "C" int aloha_synth_blobs_f32(const int32_t dims[3], int32_t num_blobs, uint64_t seed,
                                     float smin, float smax, float amin, float amax,
                                     float *out) {
  if (!dims || !out || num_blobs < 1 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_synth_blobs_f32: bad argument");
  const auto f = blob_field(dims, make_blobs(dims, num_blobs, seed, smin, smax, amin, amax));
  double mx = 0.0;
  for (double v : f) mx = std::max(mx, v);
  if (!(mx > 0.0)) mx = 1.0;
  for (size_t i = 0; i < f.size(); ++i) out[i] = (float)(f[i] / mx);
  return ALOHA_OK;
}

extern "C" int aloha_synth_fbm_f32(const int32_t dims[3], int32_t octaves, float base_period,
                                   uint64_t seed, float *out) {
  if (!dims || !out || octaves < 1 || !(base_period > 0.f) || dims[0] < 1 || dims[1] < 1 ||
      dims[2] < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_synth_fbm_f32: bad argument");
  const size_t nx = dims[0], ny = dims[1], nz = dims[2];
  SplitMix64 rng(seed);
  std::vector<uint32_t> oct_seed(octaves);
  std::vector<double> off(3 * (size_t)octaves);
  for (int o = 0; o < octaves; ++o) {
    oct_seed[o] = (uint32_t)rng.next();
    for (int a = 0; a < 3; ++a) off[3 * o + a] = rng.uniform(0.0, 1024.0);
  }
  std::vector<float> lo_t(kMaxThreads, FLT_MAX), hi_t(kMaxThreads, -FLT_MAX);
  parallel_z((int)nz, [&](unsigned t, int z0, int z1) {
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (int z = z0; z < z1; ++z)
      for (size_t y = 0; y < ny; ++y)
        for (size_t x = 0; x < nx; ++x) {
          double sum = 0.0, amp = 1.0, freq = 1.0 / base_period;
          for (int o = 0; o < octaves; ++o) {
            sum += amp * gradient_noise(x * freq + off[3 * o], y * freq + off[3 * o + 1],
                                        z * freq + off[3 * o + 2], oct_seed[o]);
            amp *= 0.5;   // gain
            freq *= 2.0;  // lacunarity
          }
          const float v = (float)sum;
          out[((size_t)z * ny + y) * nx + x] = v;
          lo = std::min(lo, v);
          hi = std::max(hi, v);
        }
    lo_t[t] = lo;
    hi_t[t] = hi;
  });
  float lo = *std::min_element(lo_t.begin(), lo_t.end());
  float hi = *std::max_element(hi_t.begin(), hi_t.end());
  const float span = hi > lo ? hi - lo : 1.f;
  parallel_z((int)nz, [&](unsigned, int z0, int z1) {
    for (size_t i = (size_t)z0 * nx * ny; i < (size_t)z1 * nx * ny; ++i)
      out[i] = std::min(1.f, std::max(0.f, (out[i] - lo) / span));
  });
  return ALOHA_OK;
}

extern "C" int aloha_synth_tf(int32_t n, int32_t num_points, uint64_t seed, float *out) {
  if (!out || n < 1 || num_points < 2)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_synth_tf: bad argument");
  SplitMix64 rng(seed);
  std::vector<double> xs(num_points);
  xs[0] = 0.0;
  xs[num_points - 1] = 1.0;
  for (int i = 1; i < num_points - 1; ++i) xs[i] = rng.uniform();
  std::sort(xs.begin() + 1, xs.end() - 1);
  std::vector<double> rgba(4 * (size_t)num_points);
  for (auto &v : rgba) v = rng.uniform();
  for (int i = 0; i < n; ++i) {
    const double u = (i + 0.5) / n;  // entry i sits at texel centre (i+.5)/n
    int k = 0;
    while (k < num_points - 2 && u > xs[k + 1]) ++k;
    const double span = xs[k + 1] - xs[k];
    const double t = span > 0 ? std::min(1.0, std::max(0.0, (u - xs[k]) / span)) : 0.0;
    for (int c = 0; c < 4; ++c)
      out[4 * i + c] = (float)(rgba[4 * k + c] + t * (rgba[4 * (k + 1) + c] - rgba[4 * k + c]));
  }
  return ALOHA_OK;
}

extern "C" int aloha_synth_camera(const aloha_box3f *bounds, int32_t W, int32_t H,
                                  float fovy_deg, float dist_factor, uint64_t seed,
                                  float out[12]) {
  if (!bounds || !out || W < 1 || H < 1 || !(fovy_deg > 0.f && fovy_deg < 180.f))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_synth_camera: bad argument");
  SplitMix64 rng(seed);
  const double az = rng.uniform(0.0, 2.0 * M_PI);
  const double el = rng.uniform(-0.5, 0.5);
  double c[3], diag2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    c[a] = 0.5 * ((double)bounds->lower[a] + bounds->upper[a]);
    const double s = (double)bounds->upper[a] - bounds->lower[a];
    diag2 += s * s;
  }
  const double dist = dist_factor * std::sqrt(diag2);
  const double dirv[3] = {std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az)};
  double org[3], fwd[3];
  for (int a = 0; a < 3; ++a) {
    org[a] = c[a] + dist * dirv[a];
    fwd[a] = -dirv[a];
  }
  const double up0[3] = {0.0, 1.0, 0.0};
  double right[3] = {fwd[1] * up0[2] - fwd[2] * up0[1], fwd[2] * up0[0] - fwd[0] * up0[2],
                     fwd[0] * up0[1] - fwd[1] * up0[0]};
  const double rl = std::sqrt(right[0] * right[0] + right[1] * right[1] + right[2] * right[2]);
  for (double &v : right) v /= rl;
  const double up[3] = {right[1] * fwd[2] - right[2] * fwd[1], right[2] * fwd[0] - right[0] * fwd[2],
                        right[0] * fwd[1] - right[1] * fwd[0]};
  const double h = 2.0 * std::tan(0.5 * fovy_deg * M_PI / 180.0);
  const double w = h * (double)W / (double)H;
  for (int a = 0; a < 3; ++a) {
    const double horiz = w * right[a], vert = h * up[a];
    out[a] = (float)org[a];
    out[3 + a] = (float)(fwd[a] - 0.5 * horiz - 0.5 * vert);  // screen lower_left
    out[6 + a] = (float)(horiz / W);
    out[9 + a] = (float)(vert / H);
  }
  return ALOHA_OK;
}
