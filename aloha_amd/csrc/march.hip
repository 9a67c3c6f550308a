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
#include "march_device.h"

namespace aloha {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr int kTileW = 16, kTileH = 16;


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
    if (kFoot || !sample_fast(b, x, y, z, s)) s = sample_exact<kFoot>(b, a.touched, db, x, y, z);
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

// One rank's spp-summed colour of one pixel: renderFrame's sample loop and
// brick do/while (deviceCode.cpp:362-393) over the bricks of rank `slot`.
// The RNG is seeded per (pixel, accumID) as the reference seeds it per launch.
template <bool kFoot>
__device__ __forceinline__ float4 rank_pixel(const MarchArgs &a, const float4 *s_tf, int slot,
                                             const float org[3], const float d[3],
                                             uint32_t pixel, float &z, uint32_t &ns) {
  const int b0 = a.rank_off[slot], b1 = a.rank_off[slot + 1];
  uint32_t rng = rng_seed(pixel, (uint32_t)a.accum_id);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int sid = 0; sid < a.spp; ++sid) {
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
        if (!slab(org, d, tmin, 1e10f, db.space_lo, db.space_hi, t0, t1)) continue;
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
        integrate<kFoot>(a, s_tf, db, org, d, bt0, bfar, j, c, ns);
        z = fminf(z, bnear / a.max_depth);
        tmin = bfar + 1e-20f;
      }
    } while (hit >= 0 && c.w < .99f);
    acc.x = acc.x + c.x;
    acc.y = acc.y + c.y;
    acc.z = acc.z + c.z;
    acc.w = acc.w + c.w;
  }
  return acc;
}

// Whether the 16x16 pixel block at (bx, by) overlaps rank slot's rectangle.
__device__ __forceinline__ bool block_meets(const MarchArgs &a, int slot, int bx, int by) {
  const int32_t *rc = a.rect[slot];
  return bx < rc[2] && bx + kTileW > rc[0] && by < rc[3] && by + kTileH > rc[1];
}

__device__ __forceinline__ bool in_rect(const MarchArgs &a, int slot, int px, int py) {
  const int32_t *rc = a.rect[slot];
  return px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3];
}

__device__ __forceinline__ void load_tf(const MarchArgs &a, float4 *s_tf) {
  for (int i = threadIdx.x; i < a.tf_n + 4; i += kBlock)
    s_tf[i] = (i >= 2 && i < a.tf_n + 2) ? a.tf[i - 2] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// island mask with the reference's gridSize formula (deviceCode.cpp:346-361):
// gridSize.x = iDivUp(pixel.x, W) is 1 for every column but the first
__device__ __forceinline__ bool island_pixel(const MarchArgs &a, int px, int py) {
  const int gx = (px + a.W - 1) / a.W;
  const int block_id = px / 16 + gx * (py / 16);
  return block_id % a.num_islands == a.island_id;
}

// Per-rank raygen: grid.z = rank slot, every pixel of every rank image is
// written.  Blocks that miss the slot's projected rectangle write the miss
// result without staging the TF or generating rays.
template <bool kFoot>
__global__ __launch_bounds__(kBlock) void march_kernel(const MarchArgs a) {
  extern __shared__ float4 s_tf[];
  const int slot = blockIdx.z;
  const int bx = blockIdx.x * kTileW, by = blockIdx.y * kTileH;
  const bool active = block_meets(a, slot, bx, by);  // block-uniform
  if (active) {
    load_tf(a, s_tf);
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = bx + (wave & 1) * 8 + (lane & 7);
  const int py = by + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const int W = a.W, H = a.H;
  const int pixel = px + W * py;
  const long long out_idx = (long long)slot * W * H + pixel;

  long long clock_begin = 0;
  if (a.flags & ALOHA_FLAG_HEATMAP) clock_begin = clock64();
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  float z = 1.f;
  uint32_t ns = 0;
  if (active && in_rect(a, slot, px, py) && island_pixel(a, px, py)) {
    float d[3];
    pixel_ray(a.dir00, a.du, a.dv, px, py, d);
    acc = rank_pixel<kFoot>(a, s_tf, slot, a.org, d, (uint32_t)pixel, z, ns);
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

// Fused raygen + sort-last composite: each pixel marches every virtual rank
// in visibility order (a.order) from transparent, exactly as that rank's
// image would hold it, and blends it 'over' the running colour with the
// composite kernel's arithmetic.  The result equals march_kernel +
// over_kernel bit for bit, without the ranks' images in HBM.  Ranks whose
// rectangle misses the pixel contribute exact zeros and are skipped.
__global__ __launch_bounds__(kBlock) void march_composite_kernel(const MarchArgs a) {
  extern __shared__ float4 s_tf[];
  const int bx = blockIdx.x * kTileW, by = blockIdx.y * kTileH;
  bool active = false;
  for (int s = 0; s < a.P; ++s) active |= block_meets(a, s, bx, by);  // block-uniform
  if (active) {
    load_tf(a, s_tf);
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = bx + (wave & 1) * 8 + (lane & 7);
  const int py = by + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const int pixel = px + a.W * py;
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t ns = 0;
  if (active && island_pixel(a, px, py)) {
    float d[3];
    pixel_ray(a.dir00, a.du, a.dv, px, py, d);
    const float inv_spp = 1.f / (float)a.spp;
    for (int k = 0; k < a.P; ++k) {
      const int slot = a.order[k];  // wave-uniform
      if (!in_rect(a, slot, px, py)) continue;
      float z = 1.f;
      const float4 acc = rank_pixel<false>(a, s_tf, slot, a.org, d, (uint32_t)pixel, z, ns);
      // rank image value (accum_id 0): (acc * 1) * (1/spp)
      const float4 c = make_float4((acc.x * 1.f) * inv_spp, (acc.y * 1.f) * inv_spp,
                                   (acc.z * 1.f) * inv_spp, (acc.w * 1.f) * inv_spp);
      const float w = 1.f - C.w;
      C.x = C.x + w * c.x;
      C.y = C.y + w * c.y;
      C.z = C.z + w * c.z;
      C.w = C.w + w * c.w;
    }
  }
  if (a.rgba) a.rgba[pixel] = C;
  if (a.rgba8)
    a.rgba8[pixel] = pack8(C.x) | (pack8(C.y) << 8) | (pack8(C.z) << 16) | (pack8(C.w) << 24);
  if (a.nsamples) a.nsamples[pixel] = ns;
}

// Batched form: V views of the same ranks in one launch.  A work item is
// (view, block tile); each pixel marches the ranks in that view's visibility
// order and blends them as march_composite_kernel does, so every view's
// image equals its own fused (and plain) frame bit for bit.  The wave tile
// shape (TW x TH pixels, TW*TH = 64) and persistence (a work counter that
// blocks pull from until it runs out) are schedule-only choices.
template <int TW, int TH>
__device__ __forceinline__ void batch_item(const BatchArgs &b, const float4 *s_tf, int item) {
  const MarchArgs &a = b.m;
  constexpr int BW = 2 * TW, BH = 2 * TH;
  const int per_view = b.tiles_x * b.tiles_y;
  const int view = item / per_view;
  const int tile = item - view * per_view;
  const int bx = (tile % b.tiles_x) * BW, by = (tile / b.tiles_x) * BH;
  const float *cam = b.cams + 12 * view;
  const int32_t *order = b.orders + a.P * view;
  const int32_t *rects = b.rects + 4 * a.P * view;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = bx + (wave & 1) * TW + (lane % TW);
  const int py = by + (wave >> 1) * TH + (lane / TW);
  if (px >= a.W || py >= a.H) return;
  const int pixel = px + a.W * py;
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t ns = 0;
  if (island_pixel(a, px, py)) {
    float d[3];
    pixel_ray(cam + 3, cam + 6, cam + 9, px, py, d);
    const float inv_spp = 1.f / (float)a.spp;
    for (int k = 0; k < a.P; ++k) {
      const int slot = order[k];  // wave-uniform
      const int32_t *rc = rects + 4 * slot;
      if (!(px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3])) continue;
      float z = 1.f;
      const float4 acc = rank_pixel<false>(a, s_tf, slot, cam, d, (uint32_t)pixel, z, ns);
      const float4 c = make_float4((acc.x * 1.f) * inv_spp, (acc.y * 1.f) * inv_spp,
                                   (acc.z * 1.f) * inv_spp, (acc.w * 1.f) * inv_spp);
      const float w = 1.f - C.w;
      C.x = C.x + w * c.x;
      C.y = C.y + w * c.y;
      C.z = C.z + w * c.z;
      C.w = C.w + w * c.w;
    }
  }
  const size_t o = (size_t)view * a.W * a.H + pixel;
  if (a.rgba) a.rgba[o] = C;
  if (a.rgba8)
    a.rgba8[o] = pack8(C.x) | (pack8(C.y) << 8) | (pack8(C.z) << 16) | (pack8(C.w) << 24);
  if (a.nsamples) a.nsamples[o] = ns;
}

template <int TW, int TH, bool kPersistent>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchArgs b) {
  extern __shared__ float4 s_tf[];
  load_tf(b.m, s_tf);
  __syncthreads();
  if (!kPersistent) {
    batch_item<TW, TH>(b, s_tf, blockIdx.x);
    return;
  }
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(b.counter, 1u);
    __syncthreads();
    const int item = s_item;
    __syncthreads();  // every wave has read s_item before it is replaced
    if (item >= b.total_items) break;
    batch_item<TW, TH>(b, s_tf, item);
  }
}

// Work-counter reset as a kernel node (graph-capturable without a memset node).
__global__ void reset_counter_kernel(uint32_t *c) { *c = 0u; }

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

hipError_t launch_march_composite(const MarchArgs &a, hipStream_t s) {
  const dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileH - 1) / kTileH, 1);
  march_composite_kernel<<<grid, kBlock, sizeof(float4) * (size_t)(a.tf_n + 4), s>>>(a);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &b, int tile_shape, int persistent_blocks,
                        hipStream_t s) {
  const size_t lds = sizeof(float4) * (size_t)(b.m.tf_n + 4);
  if (persistent_blocks > 0) {
    reset_counter_kernel<<<1, 1, 0, s>>>(b.counter);
    const dim3 grid(persistent_blocks);
    switch (tile_shape) {
      case 1: batch_kernel<16, 4, true><<<grid, kBlock, lds, s>>>(b); break;
      case 2: batch_kernel<4, 16, true><<<grid, kBlock, lds, s>>>(b); break;
      default: batch_kernel<8, 8, true><<<grid, kBlock, lds, s>>>(b); break;
    }
  } else {
    const dim3 grid(b.total_items);
    switch (tile_shape) {
      case 1: batch_kernel<16, 4, false><<<grid, kBlock, lds, s>>>(b); break;
      case 2: batch_kernel<4, 16, false><<<grid, kBlock, lds, s>>>(b); break;
      default: batch_kernel<8, 8, false><<<grid, kBlock, lds, s>>>(b); break;
    }
  }
  return hipGetLastError();
}

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
