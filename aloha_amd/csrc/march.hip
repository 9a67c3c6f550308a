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

#include <algorithm>

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
  const BrickView b = brick_view(db);
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

// integrateBrick with the samples of one ray spread over a group of G lanes:
// per step, lane gl evaluates sample k0 + gl (position, 7 taps, normal, TF,
// shade -- the costly, independent part), then every lane of the group
// applies the G samples in order with the sequential 'over' recurrence and
// the loop test (t < t1 && alpha < .99) before each.  The positions come from
// the same chain of additions t += dt, so the result is integrate()'s bit for
// bit; samples past the termination are computed and dropped.
//
// Control flow is wave-uniform wherever lanes exchange values: the step loop
// runs while any lane of the wave has work (a ballot), the shuffles execute
// on every lane, and each group's work is predicated on `run`.  All lanes of
// a group hold the same ray state, so the predicates are group-uniform.
template <int G>
__device__ __forceinline__ void group_integrate(const MarchArgs &a, const float4 *s_tf, bool run,
                                                int bid, const float o[3], const float d[3],
                                                float t0, float t1, float jitter, float4 &c,
                                                uint32_t &ns, int gl, int gbase) {
  const float dt = a.dt;
  const float segs = floorf((t0 - dt * jitter) / dt);
  float t = dt * (jitter + segs);
  if (t < t0) t += dt;
  // the brick record is read once, by the lanes that march it (inside the
  // step loop a conditional load could not be hoisted)
  BrickView b{nullptr, 0u, 0u, 0, 0, 0};
  float lo0 = 0.f, lo1 = 0.f, lo2 = 0.f;
  if (run) {
    const DevBrick &db = a.bricks[bid];
    b = brick_view(db);
    lo0 = db.vox_lo[0];
    lo1 = db.vox_lo[1];
    lo2 = db.vox_lo[2];
  }
  bool go = run && t < t1 && c.w < .99f;
  while (__any(go)) {
    float tk = t;
    for (int j = 0; j < gl; ++j) tk += dt;
    float tn = tk;  // base of the next step: G additions in all
    for (int j = gl; j < G; ++j) tn += dt;
    const bool valid = go && tk < t1;
    float4 tf = make_float4(0.f, 0.f, 0.f, 0.f);
    float shade = 0.f;
    if (valid) {
      const float x = (o[0] + tk * d[0]) - lo0;
      const float y = (o[1] + tk * d[1]) - lo1;
      const float z = (o[2] + tk * d[2]) - lo2;
      Sample s;
      if (!sample_fast(b, x, y, z, s)) s = sample_exact<false>(b, nullptr, a.bricks[bid], x, y, z);
      const float ex = s.gx + 1e-4f, ey = s.gy + 1e-4f, ez = s.gz + 1e-4f;
      const float len = sqrtf(ex * ex + ey * ey + ez * ez);
      const float nx = s.gx / len, ny = s.gy / len, nz = s.gz / len;
      tf = tf_lookup(s_tf, a.tf_n, s.val);
      shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
    }
    const int vbits = valid ? 1 : 0;
    bool live = go;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      int vj = vbits;
      float tx = tf.x, ty = tf.y, tz = tf.z, tw = tf.w, sh = shade;
      if constexpr (G > 1) {
        const int src = gbase + j;
        vj = __shfl(vbits, src);
        tx = __shfl(tf.x, src), ty = __shfl(tf.y, src), tz = __shfl(tf.z, src);
        tw = __shfl(tf.w, src), sh = __shfl(shade, src);
      }
      live = live && vj && c.w < .99f;
      if (live) {
        ++ns;
        const float w = (1.f - c.w) * tw;
        c.x = c.x + (w * tx) * sh;
        c.y = c.y + (w * ty) * sh;
        c.z = c.z + (w * tz) * sh;
        c.w = c.w + w;
      }
    }
    if (go) t = tn;
    go = go && t < t1 && c.w < .99f;
  }
}

// rank_pixel for a group of G lanes sharing pixel `pixel` (live: the group
// has a pixel to march).  Same brick loop, in wave-uniform control flow.
template <int G>
__device__ __forceinline__ float4 group_rank_pixel(const MarchArgs &a, const float4 *s_tf,
                                                   bool live, int slot, const float org[3],
                                                   const float d[3], uint32_t pixel, float &z,
                                                   uint32_t &ns, int gl, int gbase) {
  const int b0 = live ? a.rank_off[slot] : 0, b1 = live ? a.rank_off[slot + 1] : 0;
  uint32_t rng = rng_seed(pixel, (uint32_t)a.accum_id);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int sid = 0; sid < a.spp; ++sid) {
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    float tmin = 0.f;
    bool go = live;
    while (__any(go)) {
      int hit = -1;
      float best = 0.f, bnear = 0.f, bt0 = 0.f, bfar = 0.f;
      if (go) {
        for (int k = b0; k < b1; ++k) {
          const int bid = a.rank_bricks[k];
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
      }
      const bool run = go && hit >= 0;
      float j = 0.f;
      if (run) {
        bnear = bnear + tmin;
        j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(rng);
      }
      group_integrate<G>(a, s_tf, run, hit, org, d, bt0, bfar, j, c, ns, gl, gbase);
      if (run) {
        z = fminf(z, bnear / a.max_depth);
        tmin = bfar + 1e-20f;
      }
      go = run && c.w < .99f;
    }
    acc.x = acc.x + c.x;
    acc.y = acc.y + c.y;
    acc.z = acc.z + c.z;
    acc.w = acc.w + c.w;
  }
  return acc;
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

// ---- group kernels: 32x8-pixel workgroup tiles, G lanes per ray ---------
//
// A workgroup owns one 32x8 tile of one rank image (per-rank form) or of the
// final image (fused form).  Its waves pull batches of 64/G consecutive
// pixels from an LDS counter, one pixel per group of G lanes, march each
// pixel with its group (group_integrate),
// and stage the per-pixel results in LDS; then the 256 threads write the
// tile row by row (128-byte rows for the 4-byte outputs).  The per-rank grid
// lists the tiles that meet a rank's projected rectangle first, so the long
// marches start at once, and the tiles that miss every pixel of their rank
// after them: those only write the miss result.
constexpr int kGW = 32, kGH = 8;
static_assert(kGW * kGH == kBlock, "one pixel per thread in the epilogue");

struct Staged {
  float4 acc[kBlock];
  float z[kBlock];
  uint32_t ns[kBlock];
  float heat[kBlock];
  int next;
};

// tile range of rank slot s's rectangle: [tx0, tx0+ntx) x [ty0, ty0+nty)
__device__ __forceinline__ void rect_tiles(const MarchArgs &a, int s, int &tx0, int &ty0, int &ntx,
                                           int &nty) {
  const int32_t *rc = a.rect[s];
  if (rc[0] >= rc[2] || rc[1] >= rc[3]) {
    tx0 = ty0 = ntx = nty = 0;
    return;
  }
  tx0 = rc[0] / kGW;
  ty0 = rc[1] / kGH;
  ntx = (rc[2] + kGW - 1) / kGW - tx0;
  nty = (rc[3] + kGH - 1) / kGH - ty0;
}

// Per-pixel epilogue of a rank image (renderFrame :395-415), as march_kernel.
__device__ __forceinline__ void write_rank_pixel(const MarchArgs &a, long long out_idx, int px,
                                                 int py, float4 acc, float z, uint32_t ns,
                                                 float heat_t) {
  const int W = a.W, H = a.H;
  if (a.depth) a.depth[out_idx] = z;
  if (a.nsamples) a.nsamples[out_idx] = ns;
  if (a.flags & ALOHA_FLAG_HEATMAP) {
    const float3 h = heat_map(heat_t);
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

// Pixel order inside a 32x8 tile: four 8x8 sub-tiles, row-major inside
// each, so a batch of 64 (32, 16, 8) consecutive indices is an 8x8 (8x4, 8x2,
// 8x1) pixel block -- compact footprints for the lanes of a wave.
__device__ __forceinline__ void tile_pixel(int idx, int &lx, int &ly) {
  lx = (idx >> 6) * 8 + (idx & 7);
  ly = (idx >> 3) & 7;
}
__device__ __forceinline__ int tile_index(int lx, int ly) { return (lx >> 3) * 64 + ly * 8 + (lx & 7); }

// Fill part shared by the launches: `nimg` images of H rows of W pixels.
// Fill block fb of nfb takes a contiguous run of rows and calls put(img, x,
// y) for every pixel outside its image's marched box [x0, x1) x [y0, y1)
// (box(img, x0, y0, x1, y1), whole tiles, written by the march blocks);
// the 256 threads take 256 consecutive pixels per pass, so each store
// instruction covers whole 64-byte lines.  The row bookkeeping is
// block-uniform: one division per block, no per-tile index arithmetic.
template <class Box, class Put>
__device__ __forceinline__ void fill_rows(int W, int H, int nimg, int fb, int nfb, const Box &box,
                                          const Put &put) {
  const int rows = nimg * H;  // < 2^31, checked by the launchers
  const int per = (rows + nfb - 1) / nfb;
  const int r0 = per * fb, r1 = min(r0 + per, rows);
  if (r0 >= r1) return;
  int img = r0 / H, y = r0 - img * H;
  int x0, y0, x1, y1;
  box(img, x0, y0, x1, y1);
  for (int r = r0; r < r1; ++r) {
    const bool band = y >= y0 && y < y1;
    for (int x = (int)threadIdx.x; x < W; x += kBlock)
      if (!band || x < x0 || x >= x1) put(img, x, y);
    if (++y == H && r + 1 < r1) {
      y = 0;
      box(++img, x0, y0, x1, y1);
    }
  }
}

// Per-rank images.  Blocks [0, active) march the tiles that meet their rank's
// rectangle (rank by rank); the `fill_blocks` blocks after them stride over
// all (rank, tile) pairs and write the miss result of the tiles outside the
// rectangles -- few blocks, many tiles each, so the cheap part costs no
// per-block launch latency.
template <int G>
__global__ __launch_bounds__(kBlock) void march_group_kernel(const MarchArgs a, int fill_blocks) {
  extern __shared__ float4 s_tf[];
  __shared__ Staged st;
  int b = (int)blockIdx.x, slot = -1, tx = 0, ty = 0;
  for (int s = 0; s < a.P; ++s) {
    int tx0, ty0, ntx, nty;
    rect_tiles(a, s, tx0, ty0, ntx, nty);
    if (b < ntx * nty) {
      slot = s;
      tx = tx0 + b % ntx;
      ty = ty0 + b / ntx;
      break;
    }
    b -= ntx * nty;
  }
  if (slot < 0) {  // fill part (block-uniform): the pixels outside the marched tiles
    const bool heat = a.flags & ALOHA_FLAG_HEATMAP;
    const float heat_scale = a.heat_scale / (float)a.spp;
    fill_rows(
        a.W, a.H, a.P, b, fill_blocks,
        [&](int s, int &x0, int &y0, int &x1, int &y1) {
          int tx0, ty0, ntx, nty;
          rect_tiles(a, s, tx0, ty0, ntx, nty);
          x0 = tx0 * kGW, x1 = (tx0 + ntx) * kGW, y0 = ty0 * kGH, y1 = (ty0 + nty) * kGH;
        },
        [&](int s, int px, int py) {
          // a miss costs no march time: the heat map shows the clock read itself
          long long clock_begin = 0;
          if (heat) clock_begin = clock64();
          const float heat_t = heat ? (float)(clock64() - clock_begin) * heat_scale : 0.f;
          write_rank_pixel(a, (long long)s * a.W * a.H + px + (long long)a.W * py, px, py,
                           make_float4(0.f, 0.f, 0.f, 0.f), 1.f, 0u, heat_t);
        });
    return;
  }
  const int bx = tx * kGW, by = ty * kGH;
  load_tf(a, s_tf);
  if (threadIdx.x == 0) st.next = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, gl = lane % G, gbase = lane - gl;
  // one counter update per wave: its 64/G groups take consecutive pixels;
  // the batch base is wave-uniform, so the loop exit is a scalar branch.
  // Each pass takes a fresh batch, so kBlock passes always suffice.
  for (int pass = 0; pass < kBlock; ++pass) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&st.next, 64 / G);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
    if (base >= kBlock) break;
    // every lane stays in the loop body (predicated), so the group march
    // below runs in wave-uniform control flow
    const int idx = base + lane / G;
    int lx, ly;
    tile_pixel(idx, lx, ly);
    const int px = bx + lx, py = by + ly;
    long long clock_begin = 0;
    if (a.flags & ALOHA_FLAG_HEATMAP) clock_begin = clock64();
    float z = 1.f;
    uint32_t ns = 0;
    const bool live = idx < kBlock && px < a.W && py < a.H && in_rect(a, slot, px, py) &&
                      island_pixel(a, px, py);
    float d[3] = {1.f, 0.f, 0.f};
    if (live) pixel_ray(a.dir00, a.du, a.dv, px, py, d);
    const float4 acc = group_rank_pixel<G>(a, s_tf, live, slot, a.org, d,
                                           (uint32_t)(px + a.W * py), z, ns, gl, gbase);
    if (gl == 0 && idx < kBlock) {
      st.acc[idx] = acc;
      st.z[idx] = z;
      st.ns[idx] = ns;
      if (a.flags & ALOHA_FLAG_HEATMAP)
        st.heat[idx] = (float)(clock64() - clock_begin) * (a.heat_scale / (float)a.spp);
    }
  }
  __syncthreads();
  // epilogue: row-major over the tile, 128-byte rows for 4-byte outputs
  const int t = threadIdx.x;
  const int px = bx + t % kGW, py = by + t / kGW;
  if (px >= a.W || py >= a.H) return;
  const int k = tile_index(t % kGW, t / kGW);
  write_rank_pixel(a, (long long)slot * a.W * a.H + px + (long long)a.W * py, px, py, st.acc[k],
                   st.z[k], st.ns[k], st.heat[k]);
}

// Fused raygen + composite.  The work items of a tile are (rank, pixel)
// pairs of the ranks whose rectangle meets it, taken in visibility order in
// chunks of up to kFusedChunk ranks: lane groups march items in parallel and
// store each rank's pixel colour (its rank image value, accum_id 0) in LDS;
// then each thread blends its pixel's chunk colours 'over' its running
// colour in order -- march_composite's arithmetic, so the frame equals the
// rank images composited by over_kernel bit for bit.  A rank that misses a
// pixel contributes exact zeros, which 'over' leaves unchanged.
// The fused tile is 16x4 pixels (two 8x4 halves, row-major inside each), a
// quarter of the per-rank tile: its items are (rank, pixel) pairs, so the
// 256 threads still have up to kFusedChunk x 64 items per chunk while the
// grid has enough workgroups to spread the few tiles the volume covers.
constexpr int kFusedChunk = 8;
constexpr int kFW = kFusedTileW, kFH = kFusedTileH, kFPix = kFW * kFH;
static_assert(kBlock % kFPix == 0 && kFPix % 64 == 0, "items of one batch share a rank");

__device__ __forceinline__ void fused_pixel(int idx, int &lx, int &ly) {
  lx = (idx >> 5) * 8 + (idx & 7);
  ly = (idx >> 3) & 3;
}
__device__ __forceinline__ int fused_index(int lx, int ly) { return (lx >> 3) * 32 + ly * 8 + (lx & 7); }

struct FusedStage {
  float4 col[kFusedChunk][kFPix];
  uint32_t ns[kFPix];
  int ranks[ALOHA_MAX_RANKS];
  int nranks, next;
};

// Where a fused tile's frame comes from: the single frame of a launch
// (MarchArgs camera, order, rectangles) or one view of a batch (device
// tables).  Pointers into the kernel arguments are read-only constant memory.
struct FrameSrc {
  const MarchArgs &a;
  __device__ const float *org() const { return a.org; }
  __device__ const float *dir00() const { return a.dir00; }
  __device__ const float *du() const { return a.du; }
  __device__ const float *dv() const { return a.dv; }
  __device__ int order(int k) const { return a.order[k]; }
  __device__ const int32_t *rect(int s) const { return a.rect[s]; }
  __device__ long long out_base() const { return 0; }
};

struct ViewSrc {
  const BatchArgs &b;
  int view;
  __device__ const float *org() const { return b.cams + 12 * view; }
  __device__ const float *dir00() const { return b.cams + 12 * view + 3; }
  __device__ const float *du() const { return b.cams + 12 * view + 6; }
  __device__ const float *dv() const { return b.cams + 12 * view + 9; }
  __device__ int order(int k) const { return b.orders[b.m.P * view + k]; }
  __device__ const int32_t *rect(int s) const { return b.rects + 4 * (b.m.P * view + s); }
  __device__ long long out_base() const { return (long long)view * b.m.W * b.m.H; }
};

__device__ __forceinline__ bool in_rc(const int32_t *rc, int px, int py) {
  return px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3];
}

// Transparent black (no rank reaches the pixel) at frame pixel `pixel`.
__device__ __forceinline__ void fused_clear(const MarchArgs &a, long long pixel) {
  if (a.rgba) a.rgba[pixel] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (a.rgba8) a.rgba8[pixel] = 0u;
  if (a.nsamples) a.nsamples[pixel] = 0u;
}

// One fused 16x4 tile at (bx, by): rank list, chunks of (rank, pixel) items
// marched by lane groups, in-order blend, output.  The TF must be staged in
// s_tf before the call; every thread of the block calls it.  Safe to call
// repeatedly on the same `fs` when a barrier separates the calls.
template <int G, class Src>
__device__ __forceinline__ void fused_tile(const MarchArgs &a, const Src &src, FusedStage &fs,
                                           const float4 *s_tf, int bx, int by) {
  const int t = threadIdx.x;
  if (t == 0) {
    int n = 0;
    for (int k = 0; k < a.P; ++k) {
      const int s = src.order(k);
      const int32_t *rc = src.rect(s);
      if (bx < rc[2] && bx + kFW > rc[0] && by < rc[3] && by + kFH > rc[1]) fs.ranks[n++] = s;
    }
    fs.nranks = n;
  }
  if (t < kFPix) fs.ns[t] = 0u;
  __syncthreads();
  const int nranks = fs.nranks;  // block-uniform
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  const int lane = t & 63, gl = lane % G, gbase = lane - gl;
  const float inv_spp = 1.f / (float)a.spp;
  const int own = fused_index(t % kFW, (t / kFW) % kFH);  // pixel of thread t < kFPix
  for (int c0 = 0; c0 < nranks; c0 += kFusedChunk) {
    const int nc = min(kFusedChunk, nranks - c0);
    if (t == 0) fs.next = 0;
    __syncthreads();  // counter reset, previous chunk blended
    const int items = nc * kFPix;
    for (int pass = 0; pass < kFusedChunk * kFPix; ++pass) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&fs.next, 64 / G);
      base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
      if (base >= items) break;
      // 64/G consecutive items share one rank (kFPix is a multiple of 64)
      const int item = base + lane / G;
      const int r = item / kFPix, idx = item - r * kFPix;
      const int slot = fs.ranks[c0 + min(r, nc - 1)];
      int lx, ly;
      fused_pixel(idx, lx, ly);
      const int px = bx + lx, py = by + ly;
      const bool live = item < items && px < a.W && py < a.H && in_rc(src.rect(slot), px, py) &&
                        island_pixel(a, px, py);
      float d[3] = {1.f, 0.f, 0.f};
      if (live) pixel_ray(src.dir00(), src.du(), src.dv(), px, py, d);
      float z = 1.f;
      uint32_t ns = 0;
      const float4 acc = group_rank_pixel<G>(a, s_tf, live, slot, src.org(), d,
                                             (uint32_t)(px + a.W * py), z, ns, gl, gbase);
      if (gl == 0 && item < items) {
        fs.col[r][idx] = make_float4((acc.x * 1.f) * inv_spp, (acc.y * 1.f) * inv_spp,
                                     (acc.z * 1.f) * inv_spp, (acc.w * 1.f) * inv_spp);
        if (ns) atomicAdd(&fs.ns[idx], ns);
      }
    }
    __syncthreads();
    if (t < kFPix)
      for (int r = 0; r < nc; ++r) {
        const float4 c = fs.col[r][own];
        const float w = 1.f - C.w;
        C.x = C.x + w * c.x;
        C.y = C.y + w * c.y;
        C.z = C.z + w * c.z;
        C.w = C.w + w * c.w;
      }
  }
  __syncthreads();
  if (t >= kFPix) return;
  const int px = bx + t % kFW, py = by + t / kFW;
  if (px >= a.W || py >= a.H) return;
  const long long pixel = src.out_base() + px + (long long)a.W * py;
  if (a.rgba) a.rgba[pixel] = C;
  if (a.rgba8)
    a.rgba8[pixel] = pack8(C.x) | (pack8(C.y) << 8) | (pack8(C.z) << 16) | (pack8(C.w) << 24);
  if (a.nsamples) a.nsamples[pixel] = fs.ns[own];
}

// Single fused frame.  Grid: the tiles of the rectangles' bounding box
// (u = tile range x0, y0, nx, ny) first, then `fill_blocks` blocks striding
// over the other tiles of the frame.
template <int G>
__global__ __launch_bounds__(kBlock) void march_composite_group_kernel(const MarchArgs a,
                                                                       int4 u,
                                                                       int fill_blocks) {
  extern __shared__ float4 s_tf[];
  __shared__ FusedStage fs;
  const FrameSrc src{a};
  const int nact = u.z * u.w;
  if ((int)blockIdx.x >= nact) {  // fill part (block-uniform)
    fill_rows(
        a.W, a.H, 1, (int)blockIdx.x - nact, fill_blocks,
        [&](int, int &x0, int &y0, int &x1, int &y1) {
          x0 = u.x * kFW, y0 = u.y * kFH, x1 = (u.x + u.z) * kFW, y1 = (u.y + u.w) * kFH;
        },
        [&](int, int px, int py) { fused_clear(a, px + (long long)a.W * py); });
    return;
  }
  load_tf(a, s_tf);
  fused_tile<G>(a, src, fs, s_tf, (u.x + (int)blockIdx.x % u.z) * kFW,
                (u.y + (int)blockIdx.x / u.z) * kFH);
}

// Batched form: V views of the same ranks in one launch.  The active work
// items are the fused tiles inside each view's bounding box (b.boxes, prefix
// b.box_off); each is one fused_tile of that view, so every view's image
// equals its single fused frame bit for bit.  Then the blocks stride over the
// remaining tiles of every view writing transparent black.  kPersistent:
// gridDim.x blocks pull active items from a counter until it runs out;
// otherwise one block per active item plus b.fill_blocks fill blocks.
__device__ __forceinline__ void batch_active(const BatchArgs &b, int item, int &view, int &bx,
                                             int &by) {
  int lo = 0, hi = b.V;  // largest view with box_off[view] <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (b.box_off[mid] <= item) lo = mid;
    else hi = mid;
  }
  view = lo;
  const int k = item - b.box_off[view];
  const int32_t *bx4 = b.boxes + 4 * view;
  bx = (bx4[0] + k % bx4[2]) * kFW;
  by = (bx4[1] + k / bx4[2]) * kFH;
}

__device__ __forceinline__ void batch_fill(const BatchArgs &b, int fb, int nfb) {
  fill_rows(
      b.m.W, b.m.H, b.V, fb, nfb,
      [&](int view, int &x0, int &y0, int &x1, int &y1) {
        const int32_t *bx4 = b.boxes + 4 * view;
        x0 = bx4[0] * kFW, y0 = bx4[1] * kFH;
        x1 = (bx4[0] + bx4[2]) * kFW, y1 = (bx4[1] + bx4[3]) * kFH;
      },
      [&](int view, int px, int py) {
        fused_clear(b.m, (long long)view * b.m.W * b.m.H + px + (long long)b.m.W * py);
      });
}

template <int G, bool kPersistent>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchArgs b) {
  extern __shared__ float4 s_tf[];
  __shared__ FusedStage fs;
  __shared__ int s_item;
  if (!kPersistent) {
    if ((int)blockIdx.x >= b.total_active) {
      batch_fill(b, (int)blockIdx.x - b.total_active, b.fill_blocks);
      return;
    }
    load_tf(b.m, s_tf);
    int view, bx, by;
    batch_active(b, (int)blockIdx.x, view, bx, by);
    fused_tile<G>(b.m, ViewSrc{b, view}, fs, s_tf, bx, by);
    return;
  }
  load_tf(b.m, s_tf);
  for (int it = 0; it <= b.total_active; ++it) {  // bounded: each pass takes a fresh item
    if (threadIdx.x == 0) s_item = (int)atomicAdd(b.counter, 1u);
    __syncthreads();
    const int item = s_item;
    __syncthreads();  // every wave has read s_item; the previous tile's epilogue is done
    if (item >= b.total_active) break;
    int view, bx, by;
    batch_active(b, item, view, bx, by);
    fused_tile<G>(b.m, ViewSrc{b, view}, fs, s_tf, bx, by);
  }
  batch_fill(b, (int)blockIdx.x, (int)gridDim.x);
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

// fill blocks of the per-rank and fused launches: a few per CU, each
// striding over tiles
constexpr int kFillBlocks = 1024;

hipError_t launch_march_composite(const MarchArgs &a, int group, hipStream_t s) {
  // bounding box of the ranks' rectangles, in fused tiles
  int x0 = a.W, y0 = a.H, x1 = 0, y1 = 0;
  for (int r = 0; r < a.P; ++r) {
    const int32_t *rc = a.rect[r];
    if (rc[0] >= rc[2] || rc[1] >= rc[3]) continue;
    x0 = std::min(x0, (int)rc[0]), y0 = std::min(y0, (int)rc[1]);
    x1 = std::max(x1, (int)rc[2]), y1 = std::max(y1, (int)rc[3]);
  }
  int4 u = make_int4(0, 0, 0, 0);
  if (x0 < x1 && y0 < y1) {
    u.x = x0 / kFW, u.y = y0 / kFH;
    u.z = (x1 + kFW - 1) / kFW - u.x, u.w = (y1 + kFH - 1) / kFH - u.y;
  }
  const int fill_blocks = std::min<int>(a.H, kFillBlocks);
  const dim3 grid((unsigned)(u.z * u.w + fill_blocks));
  const size_t lds = sizeof(float4) * (size_t)(a.tf_n + 4);
  switch (group) {
    case 1: march_composite_group_kernel<1><<<grid, kBlock, lds, s>>>(a, u, fill_blocks); break;
    case 2: march_composite_group_kernel<2><<<grid, kBlock, lds, s>>>(a, u, fill_blocks); break;
    case 4: march_composite_group_kernel<4><<<grid, kBlock, lds, s>>>(a, u, fill_blocks); break;
    case 8: march_composite_group_kernel<8><<<grid, kBlock, lds, s>>>(a, u, fill_blocks); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &b, int lanes, int persistent_blocks, hipStream_t s) {
  const size_t lds = sizeof(float4) * (size_t)(b.m.tf_n + 4);
  if (persistent_blocks > 0) {
    reset_counter_kernel<<<1, 1, 0, s>>>(b.counter);
    const dim3 grid(persistent_blocks);
    switch (lanes) {
      case 1: batch_kernel<1, true><<<grid, kBlock, lds, s>>>(b); break;
      case 2: batch_kernel<2, true><<<grid, kBlock, lds, s>>>(b); break;
      case 4: batch_kernel<4, true><<<grid, kBlock, lds, s>>>(b); break;
      case 8: batch_kernel<8, true><<<grid, kBlock, lds, s>>>(b); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    const dim3 grid((unsigned)(b.total_active + b.fill_blocks));
    switch (lanes) {
      case 1: batch_kernel<1, false><<<grid, kBlock, lds, s>>>(b); break;
      case 2: batch_kernel<2, false><<<grid, kBlock, lds, s>>>(b); break;
      case 4: batch_kernel<4, false><<<grid, kBlock, lds, s>>>(b); break;
      case 8: batch_kernel<8, false><<<grid, kBlock, lds, s>>>(b); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_march(const MarchArgs &a, bool footprint, int group, hipStream_t s) {
  const size_t lds = sizeof(float4) * (size_t)(a.tf_n + 4);
  if (footprint) {
    const dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileH - 1) / kTileH, a.P);
    march_kernel<true><<<grid, kBlock, lds, s>>>(a);
    return hipGetLastError();
  }
  // active tiles of every rank, then the fill blocks (runs of image rows)
  const long long rows = (long long)a.P * a.H;
  if (rows > INT32_MAX) return hipErrorInvalidValue;
  const int fill_blocks = (int)std::min<long long>(rows, kFillBlocks);
  long long blocks = fill_blocks;
  for (int r = 0; r < a.P; ++r) {
    const int32_t *rc = a.rect[r];
    if (rc[0] >= rc[2] || rc[1] >= rc[3]) continue;
    const int tx0 = rc[0] / kGW, ty0 = rc[1] / kGH;
    blocks += (long long)((rc[2] + kGW - 1) / kGW - tx0) * ((rc[3] + kGH - 1) / kGH - ty0);
  }
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks);
  switch (group) {
    case 1: march_group_kernel<1><<<grid, kBlock, lds, s>>>(a, fill_blocks); break;
    case 2: march_group_kernel<2><<<grid, kBlock, lds, s>>>(a, fill_blocks); break;
    case 4: march_group_kernel<4><<<grid, kBlock, lds, s>>>(a, fill_blocks); break;
    case 8: march_group_kernel<8><<<grid, kBlock, lds, s>>>(a, fill_blocks); break;
    default: return hipErrorInvalidValue;
  }
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
