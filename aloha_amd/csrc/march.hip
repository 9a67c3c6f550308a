// march.hip -- the raygen/ray-march kernel of the per-brick render path
// (renderFrame + integrateBrick, deviceCode.cpp:340-416 and :256-298),
// written for gfx950: one 64-lane wavefront per 8x8 pixel tile, four tiles
// (16x16 pixels) per 256-thread workgroup, grid.z = virtual rank slot.
//
// Per sample the reference issues 7 trilinear texture fetches (value + 6
// central-difference taps at +-0.5 voxel) and one TF fetch.  Here the 7 taps
// read their 56 corner voxels from a 20-voxel footprint: the value tap and
// the two x taps share 12 voxels (3 x-columns of the value tap's y/z pair);
// the y and z taps each add one 2x2 row.  The x-lerps, y-lerps and z-lerps
// of every tap are still evaluated in the same order as one trilinear fetch
// (x, then y, then z, lerp(a,b,w) = fma(w, b-a, a)), so the result equals
// seven independent fetches bit for bit.  Lanes whose footprint does not
// fit the 3-wide window (a float edge at x just below 0.5, or positions
// outside the brick) take the exact per-tap path.  The TF is staged in LDS
// with two zero border entries on each side (border address mode).
//
// All arithmetic is compiled without contraction; with the explicit fma in
// lerp it matches the CPU restatement bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"

namespace aloha {
namespace {

constexpr int kBlock = 256;
constexpr int kTileW = 16, kTileH = 16;

__device__ __forceinline__ float lerp_(float a, float b, float w) {
  return __builtin_fmaf(w, b - a, a);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// --- RNG: murmur3 fmix32 seed, Numerical-Recipes LCG, 24-bit draws -------
__device__ __forceinline__ uint32_t rng_seed(uint32_t pixel, uint32_t frame) {
  uint32_t h = (pixel * 0x9E3779B1u) ^ (frame * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ float rng_next(uint32_t &s) {
  s = s * 1664525u + 1013904223u;
  return (float)(s >> 8) * (1.0f / 16777216.0f);
}

// o points at padded voxel (-1,-1,-1), so every index the taps form is
// non-negative after +1 and all offsets are 24-bit products.
struct BrickView {
  const float *o;
  uint32_t px, pxy;
  int n0, n1, n2;
};

__device__ __forceinline__ uint32_t voff(const BrickView &b, int i, int j, int k) {
  return (uint32_t)(i + 1) + __umul24((uint32_t)(j + 1), b.px) + __umul24((uint32_t)(k + 1), b.pxy);
}

// Footprint pass: mark the clamp-to-edge voxel that a corner read hits.
__device__ __forceinline__ void mark(const MarchArgs &a, const DevBrick &db, int x, int y,
                                     int z) {
  x = clampi(x, 0, db.n[0] - 1);
  y = clampi(y, 0, db.n[1] - 1);
  z = clampi(z, 0, db.n[2] - 1);
  const uint64_t idx = ((uint64_t)z * db.n[1] + y) * db.n[0] + x;
  atomicOr(&a.touched[db.touched_word + (idx >> 5)], 1u << (idx & 31));
}

// One trilinear fetch, exactly as the restated texture unit does it.
template <bool kFoot>
__device__ __forceinline__ float tap_exact(const BrickView &b, const MarchArgs &a,
                                           const DevBrick &db, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const float wx = x - fx, wy = y - fy, wz = z - fz;
  // indices stay inside the padded border; the border replicates the edge so
  // this equals per-axis clamping to [0, n-1]
  const int ix = clampi((int)fx, -1, b.n0 + 1);
  const int iy = clampi((int)fy, -1, b.n1 + 1);
  const int iz = clampi((int)fz, -1, b.n2 + 1);
  if (kFoot) {
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) mark(a, db, ix + i, iy + j, iz + k);
  }
  const float *p = b.o + voff(b, ix, iy, iz);
  const float c00 = lerp_(p[0], p[1], wx);
  const float c10 = lerp_(p[b.px], p[b.px + 1], wx);
  const float c01 = lerp_(p[b.pxy], p[b.pxy + 1], wx);
  const float c11 = lerp_(p[b.px + b.pxy], p[b.px + b.pxy + 1], wx);
  const float c0 = lerp_(c00, c10, wy);
  const float c1 = lerp_(c01, c11, wy);
  return lerp_(c0, c1, wz);
}

struct Sample {
  float val, gx, gy, gz;
};

template <bool kFoot>
__device__ __forceinline__ Sample sample_exact(const BrickView &b, const MarchArgs &a,
                                               const DevBrick &db, float x, float y, float z) {
  Sample s;
  s.gx = tap_exact<kFoot>(b, a, db, x + .5f, y, z) - tap_exact<kFoot>(b, a, db, x - .5f, y, z);
  s.gy = tap_exact<kFoot>(b, a, db, x, y + .5f, z) - tap_exact<kFoot>(b, a, db, x, y - .5f, z);
  s.gz = tap_exact<kFoot>(b, a, db, x, y, z + .5f) - tap_exact<kFoot>(b, a, db, x, y, z - .5f);
  s.val = tap_exact<kFoot>(b, a, db, x, y, z);
  return s;
}

// Shared 20-voxel footprint.  Returns false (and computes nothing) when the
// lane's taps do not fit the window; the caller then uses sample_exact.
__device__ __forceinline__ bool sample_fast(const BrickView &b, float x, float y, float z,
                                            Sample &s) {
  const float xm = x - .5f, xp = x + .5f;
  const float ym = y - .5f, yp = y + .5f;
  const float zm = z - .5f, zp = z + .5f;
  const float fxm = floorf(xm), fxc = floorf(x), fxp = floorf(xp);
  const float fym = floorf(ym), fyc = floorf(y), fyp = floorf(yp);
  const float fzm = floorf(zm), fzc = floorf(z), fzp = floorf(zp);
  const int ax = (int)fxm, cx = (int)fxc, bx = (int)fxp;
  const int ay = (int)fym, cy = (int)fyc, by = (int)fyp;
  const int az = (int)fzm, cz = (int)fzc, bz = (int)fzp;
  const int scx = cx - ax, scy = cy - ay, scz = cz - az;
  const bool ok = ((unsigned)scx <= 1u) & ((unsigned)scy <= 1u) & ((unsigned)scz <= 1u) &
                  (bx == ax + 1) & (by == ay + 1) & (bz == az + 1) & (ax >= -1) & (ax <= b.n0) &
                  (ay >= -1) & (ay <= b.n1) & (az >= -1) & (az <= b.n2);
  if (!ok) return false;
  const float wxm = xm - fxm, wxc = x - fxc, wxp = xp - fxp;
  const float wym = ym - fym, wyc = y - fyc, wyp = yp - fyp;
  const float wzm = zm - fzm, wzc = z - fzc, wzp = zp - fzp;
  const uint32_t px = b.px, pxy = b.pxy;

  // R[jj][kk][i] = V(ax+i, cy+jj, cz+kk)
  const float *r = b.o + voff(b, ax, cy, cz);
  float R[2][2][3];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int i = 0; i < 3; ++i) R[jj][kk][i] = r[i + jj * px + kk * pxy];
  // the y row outside the value tap's pair, and the z row likewise
  const int yn = scy ? ay : ay + 2;
  const int zn = scz ? az : az + 2;
  const float *ry = b.o + voff(b, cx, yn, cz);
  const float *rz = b.o + voff(b, cx, cy, zn);
  float YN[2][2], ZN[2][2];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    YN[kk][0] = ry[kk * pxy];
    YN[kk][1] = ry[kk * pxy + 1];
  }
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) {
    ZN[jj][0] = rz[jj * px];
    ZN[jj][1] = rz[jj * px + 1];
  }

  // value tap: x-lerps on the (cx, cx+1) pair
  float cX[2][2];
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const float x0 = scx ? R[jj][kk][1] : R[jj][kk][0];
      const float x1 = scx ? R[jj][kk][2] : R[jj][kk][1];
      cX[jj][kk] = lerp_(x0, x1, wxc);
    }
  const float cY0 = lerp_(cX[0][0], cX[1][0], wyc);
  const float cY1 = lerp_(cX[0][1], cX[1][1], wyc);
  s.val = lerp_(cY0, cY1, wzc);

  // x taps: pairs (ax, ax+1) and (ax+1, ax+2)
  float mX[2][2], pX[2][2];
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      mX[jj][kk] = lerp_(R[jj][kk][0], R[jj][kk][1], wxm);
      pX[jj][kk] = lerp_(R[jj][kk][1], R[jj][kk][2], wxp);
    }
  const float xm_v =
      lerp_(lerp_(mX[0][0], mX[1][0], wyc), lerp_(mX[0][1], mX[1][1], wyc), wzc);
  const float xp_v =
      lerp_(lerp_(pX[0][0], pX[1][0], wyc), lerp_(pX[0][1], pX[1][1], wyc), wzc);
  s.gx = xp_v - xm_v;

  // y taps: rows ay, ay+1, ay+2 after x-lerp with the value tap's weight
  const float yX0 = lerp_(YN[0][0], YN[0][1], wxc);  // row yn, z = cz
  const float yX1 = lerp_(YN[1][0], YN[1][1], wxc);  // row yn, z = cz+1
  // ym pair (ay, ay+1); yp pair (ay+1, ay+2)
  const float ymLo0 = scy ? yX0 : cX[0][0], ymHi0 = scy ? cX[0][0] : cX[1][0];
  const float ymLo1 = scy ? yX1 : cX[0][1], ymHi1 = scy ? cX[0][1] : cX[1][1];
  const float ypLo0 = scy ? cX[0][0] : cX[1][0], ypHi0 = scy ? cX[1][0] : yX0;
  const float ypLo1 = scy ? cX[0][1] : cX[1][1], ypHi1 = scy ? cX[1][1] : yX1;
  const float ym_v = lerp_(lerp_(ymLo0, ymHi0, wym), lerp_(ymLo1, ymHi1, wym), wzc);
  const float yp_v = lerp_(lerp_(ypLo0, ypHi0, wyp), lerp_(ypLo1, ypHi1, wyp), wzc);
  s.gy = yp_v - ym_v;

  // z taps: planes az, az+1, az+2 after x- and y-lerps with the value weights
  const float zX0 = lerp_(ZN[0][0], ZN[0][1], wxc);  // plane zn, y = cy
  const float zX1 = lerp_(ZN[1][0], ZN[1][1], wxc);  // plane zn, y = cy+1
  const float zY = lerp_(zX0, zX1, wyc);
  const float zmLo = scz ? zY : cY0, zmHi = scz ? cY0 : cY1;
  const float zpLo = scz ? cY0 : cY1, zpHi = scz ? cY1 : zY;
  s.gz = lerp_(zpLo, zpHi, wzp) - lerp_(zmLo, zmHi, wzm);
  return true;
}

// TF lookup from LDS: s_tf[k+2] = T[k], two zero entries on each side.
__device__ __forceinline__ float4 tf_lookup(const float4 *s_tf, int n, float v) {
  const float xt = v * (float)n - 0.5f;
  const float fi = floorf(xt);
  const float w = xt - fi;
  const int i = (int)fminf(fmaxf(fi, -2.f), (float)n);  // NaN -> -2 (zeros)
  const float4 e0 = s_tf[i + 2];
  const float4 e1 = s_tf[i + 3];
  return make_float4(lerp_(e0.x, e1.x, w), lerp_(e0.y, e1.y, w), lerp_(e0.z, e1.z, w),
                     lerp_(e0.w, e1.w, w));
}

__device__ __forceinline__ bool slab(const float o[3], const float d[3], float tmin, float tmax,
                                     const float lo[3], const float hi[3], float &t0,
                                     float &t1) {
  float nr[3], fr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float l = (lo[k] - o[k]) / d[k];
    const float h = (hi[k] - o[k]) / d[k];
    nr[k] = fminf(l, h);
    fr[k] = fmaxf(l, h);
  }
  t0 = fmaxf(tmin, fmaxf(fmaxf(nr[0], nr[1]), nr[2]));
  t1 = fminf(tmax, fminf(fminf(fr[0], fr[1]), fr[2]));
  return t0 < t1;
}

// integrateBrick: march one brick over [t0, t1), the interval its own
// intersect() computes from the ray's current tmin -- the same operands as
// the brick selection's slab test, so the caller passes that result.
template <bool kFoot>
__device__ __forceinline__ void integrate(const MarchArgs &a, const float4 *s_tf,
                                          const DevBrick &db, const float o[3],
                                          const float d[3], float t0, float t1, float jitter,
                                          float4 &c, uint32_t &ns) {
  const float dt = a.dt;
  const float segs = floorf((t0 - dt * jitter) / dt);
  float t = dt * (jitter + segs);
  if (t < t0) t += dt;
  const BrickView b{db.origin, (uint32_t)db.px, (uint32_t)db.pxy, db.n[0], db.n[1], db.n[2]};
  for (; t < t1 && c.w < .99f; t += dt) {
    ++ns;
    const float x = (o[0] + t * d[0]) - db.vox_lo[0];
    const float y = (o[1] + t * d[1]) - db.vox_lo[1];
    const float z = (o[2] + t * d[2]) - db.vox_lo[2];
    Sample s;
    if (kFoot || !sample_fast(b, x, y, z, s)) s = sample_exact<kFoot>(b, a, db, x, y, z);
    const float ex = s.gx + 1e-4f, ey = s.gy + 1e-4f, ez = s.gz + 1e-4f;
    const float len = sqrtf(ex * ex + ey * ey + ez * ez);
    const float nx = s.gx / len, ny = s.gy / len, nz = s.gz / len;
    const float4 tf = tf_lookup(s_tf, a.tf_n, s.val);
    const float shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
    const float w = (1.f - c.w) * tf.w;
    c.x = c.x + (w * tf.x) * shade;
    c.y = c.y + (w * tf.y) * shade;
    c.z = c.z + (w * tf.z) * shade;
    c.w = c.w + w;
  }
}

__device__ __forceinline__ float sat(float f) { return fminf(fmaxf(f, 0.f), 1.f); }

// heat-map ramp (deviceCode.cpp:300-320)
__device__ float3 heat_map(float t) {
  const float K = 4.0f / 6.0f;
  const float h = K - K * t;
  const float v = .5f + 0.5f * t;
  const float s = sat(h) * 6.0f;
  const float r = sat(fabsf(s - 3.f) - 1.0f);
  const float g = sat(2.0f - fabsf(s - 2.0f));
  const float b = sat(2.0f - fabsf(s - 4.0f));
  return make_float3(v * r, v * g, v * b);
}

__device__ __forceinline__ uint32_t pack8(float f) {
  return (uint32_t)fminf(fmaxf(f * 256.f, 0.f), 255.f);
}

template <bool kFoot>
__global__ __launch_bounds__(kBlock) void march_kernel(const MarchArgs a) {
  extern __shared__ float4 s_tf[];
  for (int i = threadIdx.x; i < a.tf_n + 4; i += kBlock)
    s_tf[i] = (i >= 2 && i < a.tf_n + 2) ? a.tf[i - 2] : make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();

  const int slot = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = blockIdx.x * kTileW + (wave & 1) * 8 + (lane & 7);
  const int py = blockIdx.y * kTileH + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const int W = a.W, H = a.H;
  const int pixel = px + W * py;
  const long long out_idx = (long long)slot * W * H + pixel;
  const int b0 = a.rank_off[slot], b1 = a.rank_off[slot + 1];

  long long clock_begin = 0;
  if (a.flags & ALOHA_FLAG_HEATMAP) clock_begin = clock64();

  // island mask with the reference's gridSize formula (deviceCode.cpp:346-361)
  const int gx = (px + W - 1) / W;  // gridSize.x: 1 for every column but the first
  const int block_id = px / 16 + gx * (py / 16);
  uint32_t rng = rng_seed((uint32_t)pixel, (uint32_t)a.accum_id);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  float z = 1.f;
  uint32_t ns = 0;
  const int32_t *rc = a.rect[slot];
  const bool may_hit = px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3];
  if (may_hit && block_id % a.num_islands == a.island_id) {
    for (int sid = 0; sid < a.spp; ++sid) {
      const float su = (float)px + .5f, sv = (float)py + .5f;
      float d[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] = (a.dir00[k] + su * a.du[k]) + sv * a.dv[k];
      const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        d[k] = d[k] / len;
        if (fabsf(d[k]) < 1e-5f) d[k] = 1e-5f;
      }
      float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
      float tmin = 0.f;
      int hit;
      do {
        // closest reported brick hit (VolBrickIsect), ties to the lower index
        hit = -1;
        float best = 0.f, bnear = 0.f, bt0 = 0.f, bfar = 0.f;
        for (int k = b0; k < b1; ++k) {
          const int bid = a.rank_bricks[k];  // k is wave-uniform: scalar loads
          const DevBrick &db = a.bricks[bid];
          float t0, t1, rep, tn;
          if (!slab(a.org, d, tmin, 1e10f, db.space_lo, db.space_hi, t0, t1)) continue;
          if (t0 > tmin) {
            rep = t0;
            tn = t0;
          } else if (t1 > tmin) {
            rep = t1;
            tn = 0.f;
          } else {
            continue;
          }
          if (hit < 0 || rep < best) {
            hit = bid;
            best = rep;
            bnear = tn;
            bt0 = t0;
            bfar = t1;
          }
        }
        if (hit >= 0) {
          bnear = bnear + tmin;
          const float j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(rng);
          const DevBrick db = a.bricks[hit];
          integrate<kFoot>(a, s_tf, db, a.org, d, bt0, bfar, j, c, ns);
          z = fminf(z, bnear / a.max_depth);
          tmin = bfar + 1e-20f;
        }
      } while (hit >= 0 && c.w < .99f);
      acc.x = acc.x + c.x;
      acc.y = acc.y + c.y;
      acc.z = acc.z + c.z;
      acc.w = acc.w + c.w;
    }
  }
  if (kFoot) return;
  if (a.depth) a.depth[out_idx] = z;
  if (a.nsamples) a.nsamples[out_idx] = ns;
  if (a.flags & ALOHA_FLAG_HEATMAP) {
    const long long clock_end = clock64();
    const float t = (float)(clock_end - clock_begin) * (a.heat_scale / (float)a.spp);
    const float3 h = heat_map(t);
    acc = make_float4(h.x + .5f * acc.x, h.y + .5f * acc.y, h.z + .5f * acc.z, .5f + .5f * acc.w);
  }
  if (a.accum) {
    if (a.accum_id > 0) {
      const float4 prev = a.accum[out_idx];
      acc.x = acc.x + prev.x;
      acc.y = acc.y + prev.y;
      acc.z = acc.z + prev.z;
      acc.w = acc.w + prev.w;
    }
    a.accum[out_idx] = acc;
  }
  const float inv_frames = 1.f / (float)(a.accum_id + 1);
  const float inv_spp = 1.f / (float)a.spp;
  const float4 fr = make_float4(acc.x * inv_frames, acc.y * inv_frames, acc.z * inv_frames,
                                acc.w * inv_frames);
  if (a.rgba)
    a.rgba[out_idx] = make_float4(fr.x * inv_spp, fr.y * inv_spp, fr.z * inv_spp, fr.w * inv_spp);
  if (a.rgba8) {
    float4 q = fr;
    if ((a.flags & ALOHA_FLAG_CROSSHAIR) && (px == W / 2 || py == H / 2))
      q = make_float4(1.f - q.x, 1.f - q.y, 1.f - q.z, 1.f - q.w);
    a.rgba8[out_idx] = pack8(q.x * inv_spp) | (pack8(q.y * inv_spp) << 8) |
                       (pack8(q.z * inv_spp) << 16) | (pack8(q.w * inv_spp) << 24);
  }
}

// Decode + pad one brick: dst element (i,j,k), i in [-1, n+2], reads the
// clamp-to-edge source voxel.  u8/u16 decode as Volume::get (u/255.f).
__global__ __launch_bounds__(256) void pad_brick_kernel(const void *src, int dtype, int nx,
                                                        int ny, int nz, float *dst, int px,
                                                        int py, long long total) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % px) - kPadLo;
    const int j = (int)((e / px) % py) - kPadLo;
    const int k = (int)(e / ((long long)px * py)) - kPadLo;
    const long long s = ((long long)clampi(k, 0, nz - 1) * ny + clampi(j, 0, ny - 1)) * nx +
                        clampi(i, 0, nx - 1);
    float v;
    if (dtype == ALOHA_U8) v = (float)((const uint8_t *)src)[s] / 255.f;
    else if (dtype == ALOHA_U16) v = (float)((const uint16_t *)src)[s] / 65535.f;
    else v = ((const float *)src)[s];
    dst[e] = v;
  }
}

}  // namespace

hipError_t launch_march(const MarchArgs &a, bool footprint, hipStream_t s) {
  const dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileH - 1) / kTileH, a.P);
  const size_t lds = sizeof(float4) * (size_t)(a.tf_n + 4);
  if (footprint)
    march_kernel<true><<<grid, kBlock, lds, s>>>(a);
  else
    march_kernel<false><<<grid, kBlock, lds, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_pad_brick(const void *src, int dtype, const int32_t n[3], float *dst_base,
                            int32_t px, int32_t py, int32_t pz, hipStream_t s) {
  const long long total = (long long)px * py * pz;
  const long long blocks = (total + 255) / 256;
  const int grid = (int)(blocks < 65536 ? blocks : 65536);
  pad_brick_kernel<<<grid, 256, 0, s>>>(src, dtype, n[0], n[1], n[2], dst_base, px, py, total);
  return hipGetLastError();
}

}  // namespace aloha
