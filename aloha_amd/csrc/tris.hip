// tris.hip -- the surface path's raygen: renderFrame with traceRay(world) and
// the ModelCH closest-hit program (deviceCode.cpp:340-416, :176-188), with
// OptiX's hardware BVH replaced by a binned-SAH BVH collapsed to four children
// per node, walked in software.
//
// Layout: one 64-lane wavefront per 8x8 pixel tile, 2x2 tiles per 256-thread
// workgroup, grid.z = rank slot (per-rank images).  Each lane keeps its
// traversal stack in LDS (4-byte entries, lane-interleaved so that the 64
// lanes of a push or pop hit 64 different banks; the per-rank launch keeps
// the first 16 there and deeper ones in private memory); a walk pushes at
// most 3 children per level, so the builder's depth bound makes 96 entries
// (and a sentinel) always enough.  A node fetch is 128 bytes (four
// children's boxes and refs) and a triangle 48 bytes.  Pixels outside the
// rectangle that a rank's geometry can project to (aloha_project_boxes of its
// root's child boxes) take the miss result without a walk.
//
// Ray/triangle test: the watertight test of Woop, Benthin and Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013 (shear to ray space, edge
// functions recomputed in double when one is exactly zero).  Closest hit is
// the smallest t > 0 below the ray's tmax (1e10, generateRay), ties to the
// smaller primitive id -- an order the BVH cannot change, so traversal gives
// exactly the brute-force closest hit of the same test.
//
// All arithmetic is compiled without contraction, so the CPU restatement
// evaluates the same operations in the same order and rounding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"
#include "march_device.h"

namespace aloha {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr int kTileW = 16, kTileH = 16;
constexpr int kLeafBits = 3;
constexpr int32_t kEmptyRoot = INT32_MIN;
constexpr int32_t kEmptyChild = 0x7fffffff;  // an unused child slot of a node
constexpr float kTMax = 1e10f;
// LDS entries of the per-rank launch's traversal stack: 16 (16 KB per
// workgroup) lets 8 waves share a SIMD instead of 5 with 32 in LDS (2.95 vs
// 3.42 ms on the 9.75 M-triangle scene with the binary BVH); the rare deeper
// entries go to private memory.  The fused launch keeps 32 in LDS.
constexpr int kRankStackLds = 16;
constexpr int kFusedStackLds = 32;
// slack on the far end of each slab interval: covers the rounding of
// (plane - org) * (1/dir) (a few ulps) so that culling stays conservative
constexpr float kSlabSlack = 1.000001f;

struct Ray {
  float o[3], d[3], inv[3];
  int kx, ky, kz;
  float sx, sy, sz;
};

__device__ __forceinline__ float pick(float x, float y, float z, int k) {
  return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ void setup_ray(const float org[3], const float d[3], Ray &r) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r.o[k] = org[k];
    r.d[k] = d[k];
    r.inv[k] = 1.f / d[k];
  }
  // dominant axis (first of equal magnitudes), then the other two in cyclic
  // order, swapped when the dominant component is negative (keeps winding)
  int kz = 0;
  if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
  if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  const float dz = pick(d[0], d[1], d[2], kz);
  if (dz < 0.f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.sx = pick(d[0], d[1], d[2], kx) / dz;
  r.sy = pick(d[0], d[1], d[2], ky) / dz;
  r.sz = 1.f / dz;
}

// t of the hit, or -1 for a miss
__device__ __forceinline__ float hit_triangle(const Ray &r, const float4 &v0, const float4 &v1,
                                              const float4 &v2) {
  const float a0 = v0.x - r.o[0], a1 = v0.y - r.o[1], a2 = v0.z - r.o[2];
  const float b0 = v1.x - r.o[0], b1 = v1.y - r.o[1], b2 = v1.z - r.o[2];
  const float c0 = v2.x - r.o[0], c1 = v2.y - r.o[1], c2 = v2.z - r.o[2];
  const float akz = pick(a0, a1, a2, r.kz), bkz = pick(b0, b1, b2, r.kz), ckz = pick(c0, c1, c2, r.kz);
  const float ax = pick(a0, a1, a2, r.kx) - r.sx * akz;
  const float ay = pick(a0, a1, a2, r.ky) - r.sy * akz;
  const float bx = pick(b0, b1, b2, r.kx) - r.sx * bkz;
  const float by = pick(b0, b1, b2, r.ky) - r.sy * bkz;
  const float cx = pick(c0, c1, c2, r.kx) - r.sx * ckz;
  const float cy = pick(c0, c1, c2, r.ky) - r.sy * ckz;
  float u = cx * by - cy * bx;
  float v = ax * cy - ay * cx;
  float w = bx * ay - by * ax;
  if (u == 0.f || v == 0.f || w == 0.f) {
    u = (float)((double)cx * (double)by - (double)cy * (double)bx);
    v = (float)((double)ax * (double)cy - (double)ay * (double)cx);
    w = (float)((double)bx * (double)ay - (double)by * (double)ax);
  }
  if ((u < 0.f || v < 0.f || w < 0.f) && (u > 0.f || v > 0.f || w > 0.f)) return -1.f;
  const float det = u + v + w;
  if (det == 0.f) return -1.f;
  const float az = r.sz * akz, bz = r.sz * bkz, cz = r.sz * ckz;
  const float tt = u * az + v * bz + w * cz;
  return tt / det;
}

// Entry and exit of the ray's [0, tmax] interval through one box.
__device__ __forceinline__ bool slab(const Ray &r, float x0, float x1, float y0, float y1,
                                     float z0, float z1, float tmax, float &tn) {
  const float ax = (x0 - r.o[0]) * r.inv[0], bx = (x1 - r.o[0]) * r.inv[0];
  const float ay = (y0 - r.o[1]) * r.inv[1], by = (y1 - r.o[1]) * r.inv[1];
  const float az = (z0 - r.o[2]) * r.inv[2], bz = (z1 - r.o[2]) * r.inv[2];
  tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.f));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
  return tn <= tf * kSlabSlack;
}

// Closest hit of one geometry below best_t (in: the limit, out: the hit t);
// returns the hit's triangle slot, or -1.  stk: this lane's LDS stack
// (stride kBlock) of kLds entries; deeper entries of a walk go to a private
// (scratch) array.  nn / nt count node fetches and triangle tests.
//
// The walk is the "while-while" loop with postponed leaves of Aila & Laine, "Understanding the Efficiency of Ray Traversal on GPUs"
// (HPG 2009): a lane that reaches a leaf keeps it and goes on through inner
// nodes until every lane of the wave still walking holds a leaf, then the
// lanes test their leaves together -- so a wave runs the costly leaf block
// once for many lanes instead of once per lane that happens to reach a leaf
// in a given trip.  The leaves are the same; the hit is the closest one with
// ties to the smaller primitive id, which no visiting order changes.
// (Culling by the best t found so far comes later, so more nodes may be
// fetched.)  The stack's bottom holds a sentinel above every node index.
constexpr int32_t kSentinel = INT32_MAX;

template <bool kStats, int kLds>
__device__ __forceinline__ int trace(const DevTris &g, const Ray &r, float &best_t,
                                        int32_t *stk, uint32_t &nn, uint32_t &nt) {
  static_assert(kLds >= 1 && kLds <= kTriStack + 1, "LDS part of the traversal stack");
  int best_slot = -1, best_prim = -1;
  if (g.root == kEmptyRoot) return -1;
  int sp = 0;
  int32_t deep[kTriStack + 1 - kLds > 0 ? kTriStack + 1 - kLds : 1];
  auto push = [&](int32_t v) {
    if (kLds == kTriStack + 1 || sp < kLds) stk[sp * kBlock] = v;
    else deep[sp - kLds] = v;
    ++sp;
  };
  auto pop = [&]() {
    --sp;
    return (kLds == kTriStack + 1 || sp < kLds) ? stk[sp * kBlock] : deep[sp - kLds];
  };
  push(kSentinel);
  int32_t node = g.root, leaf = 0;  // leaf < 0: a postponed leaf
  while (true) {
    bool searching = true;
    while ((uint32_t)node < (uint32_t)kSentinel) {  // an inner node
      const float4 *n = g.nodes + 8 * (int64_t)node;
      const float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5];
      const int4 kids = *reinterpret_cast<const int4 *>(n + 6);
      if (kStats) ++nn;
      // the children the ray meets, nearest entry first (a miss sorts last:
      // +inf); an unused slot (kEmptyChild) is a miss
      float t0, t1, t2, t3;
      const bool h0 = slab(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, best_t, t0) && kids.x != kEmptyChild;
      const bool h1 = slab(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, best_t, t1) && kids.y != kEmptyChild;
      const bool h2 = slab(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, best_t, t2) && kids.z != kEmptyChild;
      const bool h3 = slab(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, best_t, t3) && kids.w != kEmptyChild;
      float k[4] = {h0 ? t0 : INFINITY, h1 ? t1 : INFINITY, h2 ? t2 : INFINITY, h3 ? t3 : INFINITY};
      int32_t c[4] = {kids.x, kids.y, kids.z, kids.w};
      const int hits = (int)h0 + (int)h1 + (int)h2 + (int)h3;
      auto order = [&](int i, int j) {  // compare-exchange on (k, c)
        const bool sw = k[j] < k[i];
        const float tk = sw ? k[j] : k[i];
        k[j] = sw ? k[i] : k[j];
        k[i] = tk;
        const int32_t tc = sw ? c[j] : c[i];
        c[j] = sw ? c[i] : c[j];
        c[i] = tc;
      };
      order(0, 1);
      order(2, 3);
      order(0, 2);
      order(1, 3);
      order(1, 2);
      if (hits == 0) {
        node = pop();
      } else {
        if (hits > 3) push(c[3]);
        if (hits > 2) push(c[2]);
        if (hits > 1) push(c[1]);
        node = c[0];
      }
      if (node < 0 && leaf >= 0) {  // the first leaf: keep it, walk on
        searching = false;
        leaf = node;
        node = pop();
      }
      if (!__any(searching)) break;  // every lane still walking holds a leaf
    }
    if (node < 0 && leaf >= 0) {  // a leaf root
      leaf = node;
      node = pop();
    }
    while (leaf < 0) {
      const uint32_t lf = ~(uint32_t)leaf;
      const int first = (int)(lf >> kLeafBits), cnt = (int)(lf & ((1u << kLeafBits) - 1)) + 1;
      for (int i = 0; i < cnt; ++i) {
        const float4 *tp = g.tris + 3 * (int64_t)(first + i);
        const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
        if (kStats) ++nt;
        const float t = hit_triangle(r, v0, v1, v2);
        const int prim = __float_as_int(v0.w);
        if (t > 0.f && (t < best_t || (t == best_t && prim < best_prim))) {
          best_t = t;
          best_prim = prim;
          best_slot = first + i;
        }
      }
      leaf = node;  // another leaf reached as the walk broke off: it goes next
      if (node < 0) node = pop();
    }
    if (node == kSentinel) break;
  }
  return best_slot;
}

__device__ __forceinline__ float3 random_color(uint32_t idx) {
  const uint32_t r = idx * 13u * 17u + 0x234235u;
  const uint32_t g = idx * 7u * 3u * 5u + 0x773477u;
  const uint32_t b = idx * 11u * 19u + 0x223766u;
  return make_float3((float)(r & 255u) / 255.f, (float)(g & 255u) / 255.f,
                     (float)(b & 255u) / 255.f);
}

// ModelCH + the raygen's shading: geometric normal of the hit triangle,
// colour .2 + randomColor(clusterID) * max(0, -dir . n), alpha 1.
__device__ __forceinline__ float4 shade(const DevTris &g, int slot, const Ray &r) {
  const float4 *tp = g.tris + 3 * (int64_t)slot;
  const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
  const float e1x = v1.x - v0.x, e1y = v1.y - v0.y, e1z = v1.z - v0.z;
  const float e2x = v2.x - v0.x, e2y = v2.y - v0.y, e2z = v2.z - v0.z;
  const float nx = e1y * e2z - e1z * e2y;
  const float ny = e1z * e2x - e1x * e2z;
  const float nz = e1x * e2y - e1y * e2x;
  const float inv = 1.f / sqrtf(nx * nx + ny * ny + nz * nz);
  const float gx = nx * inv, gy = ny * inv, gz = nz * inv;
  const float lam = fmaxf(0.f, (-r.d[0]) * gx + (-r.d[1]) * gy + (-r.d[2]) * gz);
  const float3 cd = random_color(__float_as_uint(v1.w));
  return make_float4(.2f + cd.x * lam, .2f + cd.y * lam, .2f + cd.z * lam, 1.f);
}

__device__ __forceinline__ bool island_of(const TrisArgs &a, int px, int py) {
  const int gx = (px + a.W - 1) / a.W;  // iDivUp(launchIndex.x, launchDims.x)
  const int block_id = px / 16 + gx * (py / 16);
  return block_id % a.num_islands == a.island_id;
}

__device__ __forceinline__ void make_ray(const TrisArgs &a, int px, int py, Ray &r) {
  float d[3];
  pixel_ray(a.dir00, a.du, a.dv, px, py, d);
  setup_ray(a.org, d, r);
}

__device__ __forceinline__ uint32_t pack_rgba(const float4 &q, float s) {
  return pack8(q.x * s) | (pack8(q.y * s) << 8) | (pack8(q.z * s) << 16) | (pack8(q.w * s) << 24);
}

// Per-rank images: slot = blockIdx.z traces geometry a.geo[slot].
template <bool kStats>
__global__ __launch_bounds__(kBlock) void tris_rank_kernel(const TrisArgs a) {
  __shared__ int32_t s_stack[kRankStackLds * kBlock];
  const int slot = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = blockIdx.x * kTileW + (wave & 1) * 8 + (lane & 7);
  const int py = blockIdx.y * kTileH + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const int W = a.W, H = a.H;
  const int pixel = px + W * py;
  const long long out_idx = (long long)slot * W * H + pixel;

  long long clock_begin = 0;
  if (a.flags & ALOHA_FLAG_HEATMAP) clock_begin = clock64();
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  float z = 1.f;
  uint32_t nn = 0, nt = 0;
  const int32_t *rc = a.rect[slot];
  if (island_of(a, px, py) && px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3]) {
    Ray r;
    make_ray(a, px, py, r);
    float t = kTMax;
    const int hit = trace<kStats, kRankStackLds>(a.geo[slot], r, t, s_stack + threadIdx.x, nn, nt);
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (hit >= 0) {
      c = shade(a.geo[slot], hit, r);
      z = t / a.max_depth;
    }
    for (int sid = 0; sid < a.spp; ++sid) {
      acc.x = acc.x + c.x;
      acc.y = acc.y + c.y;
      acc.z = acc.z + c.z;
      acc.w = acc.w + c.w;
    }
  }
  if (a.depth) a.depth[out_idx] = z;
  if (kStats) {
    a.stats[2 * out_idx] = nn;
    a.stats[2 * out_idx + 1] = nt;
  }
  if (a.flags & ALOHA_FLAG_HEATMAP) {
    const unsigned long long delta = (a.flags & ALOHA_FLAG_HEAT_CLOCK)
                                         ? (unsigned long long)a.heat_clock
                                         : (unsigned long long)(clock64() - clock_begin);
    const float t = (float)delta * (a.heat_scale / (float)a.spp);
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
    a.rgba8[out_idx] = pack_rgba(q, inv_spp);
  }
}

// Fused render + depth composite.  The winner is the rank with the smallest
// z = t / maxDepth (a miss is z = 1), ties to the lower rank -- the depth
// composite's rule -- found by tracing the ranks in index order.  While the
// winner is a hit with z <= 1, later ranks only search below its t: a rank
// whose closest hit is not below it has z >= the winner's (or misses, z = 1)
// and loses the tie or the comparison either way.  The winning rank's pixel
// is then shaded exactly as its own image would hold it (accum_id 0).
template <bool kStats>
__global__ __launch_bounds__(kBlock) void tris_composite_kernel(const TrisArgs a) {
  __shared__ int32_t s_stack[kFusedStackLds * kBlock];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = blockIdx.x * kTileW + (wave & 1) * 8 + (lane & 7);
  const int py = blockIdx.y * kTileH + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const int W = a.W, H = a.H;
  const int pixel = px + W * py;
  uint32_t nn = 0, nt = 0;
  int winner = 0, win_slot = -1;
  float z_best = 1.f;
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  if (island_of(a, px, py)) {
    Ray r;
    make_ray(a, px, py, r);
    float t_win = kTMax, t_lim = kTMax;
    for (int k = 0; k < a.P; ++k) {
      float t = t_lim;
      const int32_t *rc = a.rect[k];  // outside it, rank k misses the pixel
      const bool in = px >= rc[0] && px < rc[2] && py >= rc[1] && py < rc[3];
      const int s = in ? trace<kStats, kFusedStackLds>(a.geo[k], r, t, s_stack + threadIdx.x, nn, nt) : -1;
      const float zk = s >= 0 ? t / a.max_depth : 1.f;
      if (k == 0 || zk < z_best) {
        winner = k;
        z_best = zk;
        win_slot = s;
        t_win = t;
      }
      t_lim = (win_slot >= 0 && z_best <= 1.f) ? t_win : kTMax;
    }
    if (win_slot >= 0) c = shade(a.geo[winner], win_slot, r);
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int sid = 0; sid < a.spp; ++sid) {
    acc.x = acc.x + c.x;
    acc.y = acc.y + c.y;
    acc.z = acc.z + c.z;
    acc.w = acc.w + c.w;
  }
  if (a.depth) a.depth[pixel] = z_best;
  if (kStats) {
    a.stats[2 * (long long)pixel] = nn;
    a.stats[2 * (long long)pixel + 1] = nt;
  }
  if (a.rgba8) {
    const float inv_spp = 1.f / (float)a.spp;
    float4 q = make_float4(acc.x * 1.f, acc.y * 1.f, acc.z * 1.f, acc.w * 1.f);
    if ((a.flags & ALOHA_FLAG_CROSSHAIR) && (px == W / 2 || py == H / 2))
      q = make_float4(1.f - q.x, 1.f - q.y, 1.f - q.z, 1.f - q.w);
    a.rgba8[pixel] = pack_rgba(q, inv_spp);
  }
}

}  // namespace

hipError_t launch_tris(const TrisArgs &a, bool composite, hipStream_t s) {
  const dim3 block(kBlock);
  const unsigned gx = (unsigned)((a.W + kTileW - 1) / kTileW);
  const unsigned gy = (unsigned)((a.H + kTileH - 1) / kTileH);
  if (composite) {
    if (a.stats)
      tris_composite_kernel<true><<<dim3(gx, gy, 1), block, 0, s>>>(a);
    else
      tris_composite_kernel<false><<<dim3(gx, gy, 1), block, 0, s>>>(a);
  } else {
    if (a.stats)
      tris_rank_kernel<true><<<dim3(gx, gy, (unsigned)a.P), block, 0, s>>>(a);
    else
      tris_rank_kernel<false><<<dim3(gx, gy, (unsigned)a.P), block, 0, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace aloha
