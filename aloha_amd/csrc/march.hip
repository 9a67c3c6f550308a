// march.hip -- the raygen/ray-march kernels of the per-brick render path
// (renderFrame + integrateBrick, deviceCode.cpp:340-416 and :256-298),
// written for gfx950: lane groups of 1-8 lanes per ray (march_list) over
// rays set up one per lane; the per-rank launch is a persistent grid of
// one-wave workgroups over the ranks' pixel rectangles, the fused and
// batched launches a 256-thread workgroup per 16x4 frame tile, and the
// footprint pass a 256-thread workgroup per 16x16 tile of one rank slot.
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
#include <climits>
#include <type_traits>

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
                                          const float *lut, const DevBrick &db, const float o[3],
                                          const float d[3], float t0, float t1, float jitter,
                                          float4 &c, uint32_t &ns) {
  const float dt = a.dt;
  const float segs = floorf((t0 - dt * jitter) / dt);
  float t = dt * (jitter + segs);
  if (t < t0) t += dt;
  const BrickView b = brick_view(db, lut);
  for (; t < t1 && c.w < .99f; t += dt) {
    ++ns;
    const float x = (o[0] + t * d[0]) - db.vox_lo[0];
    const float y = (o[1] + t * d[1]) - db.vox_lo[1];
    const float z = (o[2] + t * d[2]) - db.vox_lo[2];
    Sample s;
    if (kFoot || !sample_fast<true>(b, x, y, z, s))
      s = sample_exact<kFoot, true>(b, a.touched, db, x, y, z);
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

__device__ __forceinline__ void load_tf(const MarchArgs &a, float4 *s_tf) {
  for (int i = threadIdx.x; i < a.tf_n + 4; i += blockDim.x)
    s_tf[i] = (i >= 2 && i < a.tf_n + 2) ? a.tf[i - 2] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// What the march reads of one brick, cached in LDS per block (64 B): the
// brick selection and segment set-up then cost an LDS read instead of a
// dependent global load on every ray transition.
struct LdsBrick {
  float lo[3], hi[3];  // spaceRange
  float vlo[3];        // voxelRange.lower
  uint32_t px;         // padded x pitch | kLdsU8 for u8 voxels
  uint32_t pxy;
  int32_t n[3];
  const void *origin;
};
constexpr uint32_t kLdsU8 = 0x80000000u;  // pitches stay below 2^24
static_assert(sizeof(LdsBrick) == kLdsBrickBytes, "LDS brick record");

// The launch's dynamic LDS: the TF with its zero borders (tf_n + 4 float4),
// then one LdsBrick per rank_bricks entry, rank_off[0..P], the ranks' pixel
// rectangles and (a.u8_lut) the u8 decode table -- every table a lane
// indexes with a per-lane value, so that such reads are LDS reads, not
// vector loads from the kernel arguments.
struct MarchTables {
  const float4 *tf;
  const LdsBrick *brick;  // indexed like rank_bricks
  const int32_t *roff;    // rank_off[0..P]
  const int32_t *rect;    // rect[0..P) x 4
  const float *lut;       // kU8Lut entries (null when no brick stores u8)
};

// Stage the TF, the rank table and the ranks' brick records in the dynamic
// LDS `smem`; the caller synchronizes before use.  kU8 false: no brick of
// the launch stores u8 (a.u8_lut is 0), nothing of the decode compiled in.
template <bool kU8>
__device__ __forceinline__ MarchTables load_tables(const MarchArgs &a, float4 *smem) {
  load_tf(a, smem);
  LdsBrick *sb = reinterpret_cast<LdsBrick *>(smem + a.tf_n + 4);
  const int nrb = a.rank_off[a.P];
  int32_t *roff = reinterpret_cast<int32_t *>(sb + nrb);
  for (int i = threadIdx.x; i < nrb; i += blockDim.x) {
    const DevBrick &db = a.bricks[a.rank_bricks[i]];
    LdsBrick r;
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = db.space_lo[k];
      r.hi[k] = db.space_hi[k];
      r.vlo[k] = db.vox_lo[k];
      r.n[k] = db.n[k];
    }
    r.px = (uint32_t)db.px;
    if constexpr (kU8) r.px |= db.dtype == ALOHA_U8 ? kLdsU8 : 0u;
    r.pxy = (uint32_t)db.pxy;
    r.origin = db.origin;
    sb[i] = r;
  }
  int32_t *rect = roff + a.P + 1;
  for (int i = threadIdx.x; i <= a.P; i += blockDim.x) roff[i] = a.rank_off[i];
  for (int i = threadIdx.x; i < 4 * a.P; i += blockDim.x) rect[i] = a.rect[i / 4][i % 4];
  float *lut = nullptr;
  if constexpr (kU8) {
    lut = reinterpret_cast<float *>(rect + 4 * a.P);
    load_u8_lut(lut);
  }
  return MarchTables{smem, sb, roff, rect, lut};
}

// One ray of a work list: rank slot (-1: the ray can reach no brick, its
// result is the miss result), pixel, and the caller's tag for it.
struct ListRay {
  int slot, px, py, tag;
};

enum : int { kFetch = 0, kSelect = 1, kStep = 2, kDone = 3 };

// The march state of one ray: kFetch (no ray), kSelect (closest brick hit
// from tmin next), kStep (marching segment [t, t1) of brick kb), kDone.
// In a lane group every lane holds the same values.
struct RayState {
  int state, b0, b1, sid, kb;  // kb: the segment's rank_bricks index
  ListRay ray;
  uint32_t rng, ns;
  float d[3];
  float4 c, acc;  // this spp sample's colour, the spp sum
  float z, tmin, t, t1;
  long long clock_begin;
};

__device__ __forceinline__ void ray_state_init(RayState &s) {
  s.state = kFetch;
  s.b0 = s.b1 = s.sid = s.kb = 0;
  s.ray = ListRay{-1, 0, 0, 0};
  s.rng = s.ns = 0;
  s.d[0] = 1.f, s.d[1] = 0.f, s.d[2] = 0.f;
  s.c = s.acc = make_float4(0.f, 0.f, 0.f, 0.f);
  s.z = 1.f, s.tmin = 0.f, s.t = 0.f, s.t1 = 0.f;
  s.clock_begin = 0;
}

// End of one spp sample: sum it, start the next or hand the ray back
// (store(ray, acc, z, ns, clock_begin), called where `owner`).
template <class Store>
__device__ __forceinline__ void end_sample(const MarchArgs &a, RayState &s, bool owner,
                                           const Store &store) {
  s.acc.x = s.acc.x + s.c.x;
  s.acc.y = s.acc.y + s.c.y;
  s.acc.z = s.acc.z + s.c.z;
  s.acc.w = s.acc.w + s.c.w;
  if (++s.sid < a.spp) {
    s.c = make_float4(0.f, 0.f, 0.f, 0.f);
    s.tmin = 0.f;
    s.state = kSelect;
  } else {
    if (owner) store(s.ray, s.acc, s.z, s.ns, s.clock_begin);
    s.state = kFetch;
  }
}

// kSelect: the closest reported brick hit from tmin (VolBrickIsect, ties to
// the lower index; deviceCode.cpp:366-389): sets up the segment (firstSampleT,
// depth, next tmin) or ends the spp sample.  Returns whether a brick was hit.
template <class Store>
__device__ __forceinline__ bool select_brick(const MarchArgs &a, const MarchTables &tb,
                                             const float org[3], RayState &s, bool owner,
                                             const Store &store) {
  const float dt = a.dt;
  int hit = -1;
  float best = 0.f, bnear = 0.f, bt0 = 0.f, bfar = 0.f;
  for (int kb = s.b0; kb < s.b1; ++kb) {
    const LdsBrick &db = tb.brick[kb];
    float s0, s1, rep, tn;
    if (!slab(org, s.d, s.tmin, 1e10f, db.lo, db.hi, s0, s1)) continue;
    if (s0 > s.tmin) {
      rep = s0;
      tn = s0;
    } else if (s1 > s.tmin) {
      rep = s1;
      tn = 0.f;
    } else {
      continue;
    }
    if (hit < 0 || rep < best) {  // ties keep the earlier entry
      hit = kb;
      best = rep;
      bnear = tn;
      bt0 = s0;
      bfar = s1;
    }
  }
  if (hit < 0) {
    end_sample(a, s, owner, store);
    return false;
  }
  bnear = bnear + s.tmin;
  const float j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(s.rng);
  // firstSampleT
  const float segs = floorf((bt0 - dt * j) / dt);
  float t = dt * (j + segs);
  if (t < bt0) t += dt;
  s.t = t;
  s.t1 = bfar;
  s.kb = hit;
  s.z = fminf(s.z, bnear / a.max_depth);
  s.tmin = bfar + 1e-20f;
  if (t < s.t1 && s.c.w < .99f) s.state = kStep;
  else if (!(s.c.w < .99f)) end_sample(a, s, owner, store);  // else: the next brick next
  return true;
}

// Take ray r: generateRay at the pixel centre, the RNG seeded per (pixel,
// accumID), then the first brick selection.
template <class Store>
__device__ __forceinline__ bool begin_ray(const MarchArgs &a, const MarchTables &tb,
                                          const float org[3], const float dir00[3],
                                          const float du[3], const float dv[3], RayState &s,
                                          const ListRay &r, bool owner, const Store &store) {
  s.ray = r;
  // slot -1: a pixel no brick can reach (no brick to select: the miss result)
  s.b0 = r.slot < 0 ? 0 : tb.roff[r.slot];
  s.b1 = r.slot < 0 ? 0 : tb.roff[r.slot + 1];
  pixel_ray(dir00, du, dv, r.px, r.py, s.d);
  if (a.jitter_mode != ALOHA_JITTER_FIXED)
    s.rng = rng_seed((uint32_t)(r.px + a.W * r.py), (uint32_t)a.accum_id);
  s.sid = 0;
  s.ns = 0;
  s.z = 1.f;
  s.acc = make_float4(0.f, 0.f, 0.f, 0.f);
  s.c = s.acc;
  s.tmin = 0.f;
  s.state = kSelect;
  return select_brick(a, tb, org, s, owner, store);
}

// A ray after begin_ray, as one lane set it up for a lane group to march
// (48 B, three 16-byte LDS words).  ns, c and acc are zero at that point:
// the first selection takes no sample, and a sample that ends there sums
// zeros.  tmin is not stored: after a hit it is t1 + 1e-20 (t1 = the brick
// exit), after a miss (which ends the first spp sample, so sid > 0) it is 0.
// Brick indices fit 9 bits (ALOHA_MAX_RANK_BRICKS <= 256).
struct RayRec {
  int4 w0;    // px, py, tag, kb | b0 << 9 | b1 << 18
  float4 w1;  // d[0..2], rng (bits)
  float4 w2;  // z, t, t1, state | sid << 2 (bits)
};
static_assert(sizeof(RayRec) == kRayRecBytes, "LDS ray record");
static_assert(ALOHA_MAX_RANK_BRICKS <= 256, "brick indices of a ray record");

__device__ __forceinline__ void rec_pack(RayRec &r, const RayState &s) {
  r.w0 = make_int4(s.ray.px, s.ray.py, s.ray.tag, s.kb | (s.b0 << 9) | (s.b1 << 18));
  r.w1 = make_float4(s.d[0], s.d[1], s.d[2], __uint_as_float(s.rng));
  r.w2 = make_float4(s.z, s.t, s.t1, __int_as_float(s.state | (s.sid << 2)));
}

__device__ __forceinline__ void rec_adopt(RayState &s, const RayRec &r) {
  s.ray = ListRay{0, r.w0.x, r.w0.y, r.w0.z};  // slot: only begin_ray reads it
  s.kb = r.w0.w & 0x1FF, s.b0 = (r.w0.w >> 9) & 0x1FF, s.b1 = r.w0.w >> 18;
  s.d[0] = r.w1.x, s.d[1] = r.w1.y, s.d[2] = r.w1.z;
  s.rng = __float_as_uint(r.w1.w);
  s.z = r.w2.x, s.t = r.w2.y, s.t1 = r.w2.z;
  const int ss = __float_as_int(r.w2.w);
  s.state = ss & 3, s.sid = ss >> 2;
  s.tmin = s.sid > 0 ? 0.f : s.t1 + 1e-20f;
  s.ns = 0;
  s.c = s.acc = make_float4(0.f, 0.f, 0.f, 0.f);
}

// renderFrame's per-pixel work (sample loop, brick do/while, integrateBrick;
// deviceCode.cpp:362-393 and :256-298) for the rays of a work list [0,
// count), marched by lane groups of G lanes with persistent-lane scheduling:
// a group takes the next list entry as soon as its ray is done, so groups
// stay busy while entries remain instead of waiting for the longest ray of a
// fixed batch.
//
// Each group runs the RayState machine, one transition or one march step per
// loop trip.  In a step lane gl evaluates sample k0 + gl (position, 7 taps,
// normal, TF, shade -- the costly, independent part); then every lane of the
// group applies the G samples in order with the sequential 'over' recurrence
// and the loop test (t < t1 && alpha < .99) before each.  Positions come from
// the same chain of additions t += dt, the segment start from firstSampleT,
// the jitter from the same RNG draws in the same order, so every ray's result
// equals rank_pixel()'s bit for bit; samples past the termination are
// computed and dropped.  A segment that ends in a rank of one brick ends the
// spp sample at once: the next selection could only test that brick again,
// and from tmin = its exit (+1e-20) its slab interval [max(tmin, near), exit)
// is empty, so that selection finds no hit.
//
// Control flow is wave-uniform wherever lanes exchange values: the loop runs
// while any group of the wave has a ray (a ballot), the broadcasts execute on
// every lane, and each group's work is predicated on its state, which is the
// same on all of its lanes.  fetch(want) is called by every lane of the
// wave and returns, on the lanes with want set (the first lane of each group
// without a ray), the next list index, a negative value for "none yet, ask
// again next trip" or count and above when the list is exhausted.  entry(k)
// gives list entry k: a ListRay (kRec false: the group sets the ray up) or a
// RayRec (kRec true: set up already).  store(ray, acc, z, ns, clock_begin) is
// called on the group's first lane when its ray is done.
template <int G, bool kRec, bool kU8, class Fetch, class Entry, class Store>
__device__ __forceinline__ void march_list(const MarchArgs &a, const MarchTables &tb,
                                           const float org[3], const float dir00[3],
                                           const float du[3], const float dv[3],
                                           const Fetch &fetch, int count, const Entry &entry,
                                           const Store &store) {
  const int lane = threadIdx.x & 63, gl = lane % G, gbase = lane - gl;
  const bool owner = gl == 0;
  const float dt = a.dt;
  RayState s;
  ray_state_init(s);
  BrickView bv{nullptr, 0u, 0u, 0, 0, 0, nullptr, false};
  float lo0 = 0.f, lo1 = 0.f, lo2 = 0.f;
  bool fresh = false;  // a new segment: its brick view is still to be read
  // One trip: a group whose segment ended in a rank of several bricks
  // selects its next brick (or ends its ray), a group without a ray takes
  // the next entry (and selects its first brick), then every group with a
  // segment steps -- so a ray's transitions cost no extra trips.
  while (__any(s.state != kDone)) {
    if (s.state == kSelect) fresh |= select_brick(a, tb, org, s, owner, store);
    int k = fetch(s.state == kFetch && owner);
    k = group_bcast_i<G, 0>(k, gbase);
    if (s.state == kFetch) {
      if (k >= count) {
        s.state = kDone;
      } else if (k >= 0) {
        if constexpr (kRec) {
          rec_adopt(s, entry(k));
          fresh = true;
        } else {
          fresh |= begin_ray(a, tb, org, dir00, du, dv, s, entry(k), owner, store);
        }
        if (a.flags & ALOHA_FLAG_HEATMAP) s.clock_begin = clock64();
      }
    }
    // kStep: G samples of the group's segment
    const bool go = s.state == kStep;
    if (go && fresh) {
      const LdsBrick &db = tb.brick[s.kb];
      // kU8 false: no brick of the launch stores u8, and the decode path
      // is not compiled in
      if constexpr (kU8)
        bv = BrickView{(const gfloat *)db.origin, db.px & ~kLdsU8, db.pxy, db.n[0], db.n[1],
                       db.n[2], tb.lut, (db.px & kLdsU8) != 0u};
      else  // no u8 brick in the launch: no flag in the pitch
        bv = BrickView{(const gfloat *)db.origin, db.px, db.pxy, db.n[0], db.n[1], db.n[2],
                       nullptr, false};
      lo0 = db.vlo[0];
      lo1 = db.vlo[1];
      lo2 = db.vlo[2];
      fresh = false;
    }
    // the chain t, t + dt, (t + dt) + dt, ...: this lane's sample is element
    // gl, the next step's base element G
    float tk = s.t, tn = s.t;
#pragma unroll
    for (int j = 1; j <= G; ++j) {
      tn += dt;
      if (j == gl) tk = tn;
    }
    const bool valid = go && tk < s.t1;
    float4 tf = make_float4(0.f, 0.f, 0.f, 0.f);
    // an evaluated sample has shade >= 0.1 or NaN, so shade < 0 marks "no
    // sample" (one value less to broadcast)
    float shade = -1.f;
    if (valid) {
      const float *d = s.d;
      const float x = (org[0] + tk * d[0]) - lo0;
      const float y = (org[1] + tk * d[1]) - lo1;
      const float zz = (org[2] + tk * d[2]) - lo2;
      Sample sm;
      if (!sample_fast<kU8>(bv, x, y, zz, sm))
        sm = sample_exact<false, kU8>(bv, nullptr, a.bricks[0], x, y, zz);  // no footprint: db unread
      const float ex = sm.gx + 1e-4f, ey = sm.gy + 1e-4f, ez = sm.gz + 1e-4f;
      const float len = sqrtf(ex * ex + ey * ey + ez * ez);
      const float nx = sm.gx / len, ny = sm.gy / len, nz = sm.gz / len;
      tf = tf_lookup(tb.tf, a.tf_n, sm.val);
      shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
    }
    bool live = go;
    static_for<0, G>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      const float tx = group_bcast<G, j>(tf.x, gbase), ty = group_bcast<G, j>(tf.y, gbase);
      const float tz = group_bcast<G, j>(tf.z, gbase), tw = group_bcast<G, j>(tf.w, gbase);
      const float sh = group_bcast<G, j>(shade, gbase);
      live = live && !(sh < 0.f) && s.c.w < .99f;
      if (live) {
        ++s.ns;
        const float w = (1.f - s.c.w) * tw;
        s.c.x = s.c.x + (w * tx) * sh;
        s.c.y = s.c.y + (w * ty) * sh;
        s.c.z = s.c.z + (w * tz) * sh;
        s.c.w = s.c.w + w;
      }
    });
    if (go) {
      s.t = tn;
      if (!(s.t < s.t1 && s.c.w < .99f)) {  // segment over
        if (s.c.w < .99f && s.b1 - s.b0 > 1) s.state = kSelect;
        else end_sample(a, s, owner, store);
      }
    }
  }
}

// One rank's spp-summed colour of one pixel: renderFrame's sample loop and
// brick do/while (deviceCode.cpp:362-393) over the bricks of rank `slot`.
// The RNG is seeded per (pixel, accumID) as the reference seeds it per launch.
template <bool kFoot>
__device__ __forceinline__ float4 rank_pixel(const MarchArgs &a, const float4 *s_tf,
                                             const float *lut, int slot,
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
        integrate<kFoot>(a, s_tf, lut, db, org, d, bt0, bfar, j, c, ns);
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

// island mask with the reference's gridSize formula (deviceCode.cpp:346-361):
// gridSize.x = iDivUp(pixel.x, W) is 1 for every column but the first
__device__ __forceinline__ bool island_pixel(const MarchArgs &a, int px, int py) {
  if (a.num_islands == 1) return a.island_id == 0;  // block_id % 1 == 0
  const int gx = px > 0 ? 1 : 0;  // (px + W - 1) / W for 0 <= px < W
  const int block_id = px / 16 + gx * (py / 16);
  return block_id % a.num_islands == a.island_id;
}

// Footprint pass: the per-lane march of every rank pixel (grid.z = rank
// slot) with each corner read marking its voxel in the `touched` bitset;
// blocks that miss the slot's projected rectangle do nothing.
template <bool kFoot>
__global__ __launch_bounds__(kBlock) void march_kernel(const MarchArgs a) {
  static_assert(kFoot, "the footprint pass is the only per-lane march launch");
  extern __shared__ float4 s_tf[];
  const int slot = blockIdx.z;
  const int bx = blockIdx.x * kTileW, by = blockIdx.y * kTileH;
  const bool active = block_meets(a, slot, bx, by);  // block-uniform
  // the u8 decode table, staged whatever the bricks' storage: the samplers
  // compiled with the decode read it on every lane
  float *lut = reinterpret_cast<float *>(s_tf + a.tf_n + 4);
  if (active) {
    load_tf(a, s_tf);
    load_u8_lut(lut);
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = bx + (wave & 1) * 8 + (lane & 7);
  const int py = by + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  float z = 1.f;
  uint32_t ns = 0;
  if (active && in_rect(a, slot, px, py) && island_pixel(a, px, py)) {
    float d[3];
    pixel_ray(a.dir00, a.du, a.dv, px, py, d);
    (void)rank_pixel<kFoot>(a, s_tf, lut, slot, a.org, d, (uint32_t)(px + a.W * py), z, ns);
  }
}

// Lane-parallel ray set-up with stream compaction: every thread of the block
// runs setup(i, state) for items i in [0, n) (a begin_ray, or nothing for an
// item without a ray); the rays still marching (state other than kFetch) go
// to recs as RayRec records, in item order.  Returns the count,
// block-uniform.  Every thread of the block calls it; it ends with a
// barrier, so LDS written before the call is visible after it.
template <class Setup>
__device__ __forceinline__ int compact_records(int n, RayRec *recs, int *wcnt, const Setup &setup) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int total = 0;
  for (int i0 = 0; i0 < n; i0 += kBlock) {
    const int i = i0 + t;
    RayState st;
    ray_state_init(st);
    if (i < n) setup(i, st);
    const bool on = st.state != kFetch;
    const uint64_t m = __ballot(on);
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    int base = total, sum = total;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) {
      base += k < w ? wcnt[k] : 0;
      sum += wcnt[k];
    }
    if (on) rec_pack(recs[base + __popcll(m & ((1ull << lane) - 1ull))], st);
    __syncthreads();  // wcnt is rewritten by the next round; the records are complete
    total = sum;
  }
  return total;
}

// heatMap's argument (deviceCode.cpp:399-401): the pixel's clock delta times
// heatMapScale / spp (`scale`), the delta measured since clock_begin or,
// under ALOHA_FLAG_HEAT_CLOCK, the host-given a.heat_clock.
__device__ __forceinline__ float heat_time(const MarchArgs &a, long long clock_begin, float scale) {
  const unsigned long long delta = (a.flags & ALOHA_FLAG_HEAT_CLOCK)
                                       ? (unsigned long long)a.heat_clock
                                       : (unsigned long long)(clock64() - clock_begin);
  return (float)delta * scale;
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
  const float inv_frames = a.inv_frames, inv_spp = a.inv_spp;
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

// ---------------------------------------------------------------------------
// Persistent per-rank launch.
//
// The work of a frame is every pixel of every rank's rectangle (the pixels a
// rank's bricks can cover) plus the miss result everywhere else.  The grid
// is one wave per block, as many as the GPU holds at once.  The rectangle
// pixels are cut into runs of up to kRankRun consecutive pixels of one row,
// numbered rank after rank and row-major (q.run_off); the run list is cut
// into kRankParts contiguous parts, and the waves w = x (mod kRankParts) --
// all on one XCD, as workgroups are dealt to the XCDs in turn -- take part
// x's runs round-robin.  So every wave gets the same number of runs (within
// one), its runs are spread over the part (no wave holds a whole expensive
// region), the rays of a run are neighbours (their voxel reads share cache
// lines), and the voxels one XCD reads come from one part of the frame.
// Static shares need no atomics and no counters to reset between launches.
//
// The wave takes its runs 64 / kRankRun at a time.  Every lane sets up the
// ray of one pixel (generateRay, RNG seed, first brick selection: the
// per-ray work, here at one ray per lane instead of one per lane group); a
// pixel whose ray reaches no brick, or fails the island test, stores its
// result right there, and the other rays go to a 64-entry LDS list as RayRec
// records, in pixel order.  Lane groups then march the list with
// march_list; when it is used up the wave sets up its next runs.  A wave
// that is done writes the miss result of its share of the image rows outside
// the rectangles.
__device__ __forceinline__ long long rank_index(const MarchArgs &a, int slot, int px, int py) {
  return (long long)slot * a.W * a.H + px + (long long)a.W * py;
}

template <int G, bool kU8>
__global__ __launch_bounds__(64) void march_persist_kernel(const MarchArgs a, const RanksArgs q) {
  static_assert(64 % kRankRun == 0, "a set-up pass covers whole runs");
  constexpr int kRunsPerPass = 64 / kRankRun;
  extern __shared__ float4 smem[];
  const MarchTables tb = load_tables<kU8>(a, smem);
  RayRec *recs = reinterpret_cast<RayRec *>(smem + march_tables_bytes(a) / sizeof(float4) + 1);
  int32_t *roff = reinterpret_cast<int32_t *>(recs + 64);
  int32_t *cbase = roff + a.P + 1;
  const int lane = (int)threadIdx.x, w = (int)blockIdx.x, nw = (int)gridDim.x;
  for (int i = lane; i <= a.P; i += 64) roff[i] = q.run_off[i];
  if (q.compact)
    for (int i = lane; i < a.P; i += 64) cbase[i] = q.cbase[i];
  __syncthreads();
  const int runs = q.run_off[a.P];
  // this wave's part [r0, r1) of the run list and its place among the
  // part's nwp waves
  const int parts = nw < kRankParts ? 1 : kRankParts;
  const int part = w % parts, wi = w / parts, nwp = (nw - part + parts - 1) / parts;
  const int r0 = (int)((long long)runs * part / parts);
  const int r1 = (int)((long long)runs * (part + 1) / parts);
  const int my_runs = r1 - r0 > wi ? (r1 - r0 - wi + nwp - 1) / nwp : 0;
  const int passes = (my_runs + kRunsPerPass - 1) / kRunsPerPass;
  const float heat_scale = a.heat_scale / (float)a.spp;
  auto store = [&](const ListRay &r, float4 acc, float z, uint32_t ns, long long clock_begin) {
    if (r.tag < 0) return;
    const float heat_t = (a.flags & ALOHA_FLAG_HEATMAP) ? heat_time(a, clock_begin, heat_scale) : 0.f;
    long long o;
    if (q.compact) {  // the pixel's place in its slot's rectangle
      const int32_t *rc = tb.rect + 4 * r.tag;
      o = cbase[r.tag] + (long long)(r.py - rc[1]) * (rc[2] - rc[0]) + (r.px - rc[0]);
    } else {
      o = rank_index(a, r.tag, r.px, r.py);
    }
    write_rank_pixel(a, o, r.px, r.py, acc, z, ns, heat_t);
  };
  // set-up pass p: lane l takes pixel l % kRankRun of the wave's run
  // kRunsPerPass p + l / kRankRun; returns the number of records (wave-uniform)
  auto setup = [&](int p) {
    const int m = kRunsPerPass * p + lane / kRankRun;  // the wave's m-th run
    const int run = r0 + wi + m * nwp;
    RayState s;
    ray_state_init(s);
    if (m < my_runs) {
      int slot = 0;  // the rank whose runs hold this one: run_off[slot] <= run < run_off[slot + 1]
      while (roff[slot + 1] <= run) ++slot;
      const int32_t *rc = tb.rect + 4 * slot;
      const int l = run - roff[slot], cpr = (rc[2] - rc[0] + kRankRun - 1) / kRankRun;
      const int row = l / cpr;
      const int px = rc[0] + (l - row * cpr) * kRankRun + lane % kRankRun, py = rc[1] + row;
      if (px < rc[2]) {
        if (a.flags & ALOHA_FLAG_HEATMAP) s.clock_begin = clock64();
        // tag: the rank slot (the slot is -1 when no brick can be reached)
        const ListRay r{island_pixel(a, px, py) ? slot : -1, px, py, slot};
        begin_ray(a, tb, a.org, a.dir00, a.du, a.dv, s, r, true, store);
      }
    }
    const bool live = s.state != kFetch;  // kFetch: no pixel, or its result is written
    const uint64_t mk = __ballot(live);
    if (live) rec_pack(recs[__popcll(mk & ((1ull << lane) - 1ull))], s);
    return (int)__popcll(mk);
  };
  int pass = 0, nrec = 0, next = 0;  // wave-uniform
  auto fetch = [&](bool want) {
    if (__any(want) && next >= nrec && pass < passes) {
      __syncthreads();  // every record of the list has been read
      do {
        nrec = setup(pass++);
      } while (nrec == 0 && pass < passes);
      next = 0;
      __syncthreads();  // the new list is written
    }
    const uint64_t m = __ballot(want);
    const int k = next + __popcll(m & ((1ull << lane) - 1ull));
    next += __popcll(m);
    return k < nrec ? k : pass < passes ? -1 : INT_MAX;
  };
  march_list<G, true, kU8>(a, tb, a.org, a.dir00, a.du, a.dv, fetch, INT_MAX,
                      [&](int k) { return recs[k]; }, store);

  // fill: the miss result of every pixel outside the rectangles, rows
  // w, w + nw, ... of the P x H rank image rows, then (for a frame) the H
  // rows of the composite outputs outside the composite's box
  const bool heat = a.flags & ALOHA_FLAG_HEATMAP;
  const bool plain = !(a.flags & (ALOHA_FLAG_HEATMAP | ALOHA_FLAG_CROSSHAIR)) &&
                     !(a.accum && a.accum_id > 0);
  const int rank_rows = q.compact ? 0 : a.P * a.H;  // compact images: nothing outside
  const int rows = rank_rows + (q.frame || q.frame8 ? a.H : 0);
  for (int r = w; r < rows; r += nw) {
    if (r >= rank_rows) {
      const int y = r - rank_rows;
      const bool band = y >= q.box.y && y < q.box.w;
      for (int x = lane; x < a.W; x += 64) {
        if (band && x >= q.box.x && x < q.box.z) continue;  // composited
        const long long o = x + (long long)a.W * y;
        if (q.frame) store_nt(q.frame + o, make_float4(0.f, 0.f, 0.f, 0.f));
        if (q.frame8) store_nt(q.frame8 + o, 0u);
      }
      continue;
    }
    const int s = r / a.H, y = r - s * a.H;
    const int32_t *rc = tb.rect + 4 * s;
    const bool band = rc[0] < rc[2] && y >= rc[1] && y < rc[3];
    for (int x = lane; x < a.W; x += 64) {
      if (band && x >= rc[0] && x < rc[2]) continue;  // a rectangle pixel: marched
      const long long o = rank_index(a, s, x, y);
      if (plain) {  // streaming stores: they overlap the march of the other waves
        const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
        if (a.depth) store_nt(a.depth + o, 1.f);
        if (a.nsamples) store_nt(a.nsamples + o, 0u);
        if (a.accum) store_nt(a.accum + o, zero);
        if (a.rgba) store_nt(a.rgba + o, zero);
        if (a.rgba8) store_nt(a.rgba8 + o, 0u);
      } else {
        long long clock_begin = 0;
        if (heat) clock_begin = clock64();
        const float heat_t = heat ? heat_time(a, clock_begin, heat_scale) : 0.f;
        write_rank_pixel(a, o, x, y, make_float4(0.f, 0.f, 0.f, 0.f), 1.f, 0u, heat_t);
      }
    }
  }
}

// Fused raygen + composite.  The work items of a tile are (rank, pixel)
// pairs of the ranks whose rectangle meets it, taken in visibility order in
// chunks of up to kFusedChunk ranks: every thread sets up the rays of its
// items (compact_records), lane groups march the list in parallel and store
// each rank's pixel colour (its rank image value, accum_id 0) in LDS;
// then each thread blends its pixel's chunk colours 'over' its running
// colour in order -- march_composite's arithmetic, so the frame equals the
// rank images composited by over_kernel bit for bit.  A rank that misses a
// pixel contributes exact zeros, which 'over' leaves unchanged.
// The fused tile is 16x4 pixels (two 8x4 halves, row-major inside each), a
// quarter of the per-rank tile: its items are (rank, pixel) pairs, so the
// 256 threads still have up to kFusedChunk x 64 items per chunk while the
// grid has enough workgroups to spread the few tiles the volume covers.  Six
// ranks per chunk keep the workgroup's LDS (the chunk's ray records and
// colours, ~26 KB, plus the brick tables) small enough for four workgroups
// per CU with 64 bricks.
constexpr int kFusedChunk = 6;
constexpr int kFW = kFusedTileW, kFH = kFusedTileH, kFPix = kFW * kFH;
static_assert(kBlock % kFPix == 0 && kFPix % 64 == 0, "items of one batch share a rank");

// A fused tile of 64 pixels, 8 << (3 - hl) wide and 1 << hl high (8x8, 16x4
// or 32x2): item idx is pixel (lx, ly) of its 8-wide column block, row-major
// inside the block, so the 64 items of a wave are an 8 x (1 << hl) patch.
struct TileShape {
  int wl, hl;  // log2 of width and height
  __device__ int w() const { return 1 << wl; }
  __device__ int h() const { return 1 << hl; }
  __device__ void pixel(int idx, int &lx, int &ly) const {
    lx = (idx >> (3 + hl)) * 8 + (idx & 7);
    ly = (idx >> 3) & ((1 << hl) - 1);
  }
  __device__ int index(int lx, int ly) const { return ((lx >> 3) << (3 + hl)) | (ly << 3) | (lx & 7); }
};
constexpr int kFusedWl = 4, kFusedHl = 2;  // the single fused frame's 16x4
static_assert((1 << kFusedWl) == kFW && (1 << kFusedHl) == kFH, "fused tile shape");

struct FusedStage {
  float4 col[kFusedChunk][kFPix];
  uint32_t ns[kFPix];
  int ranks[ALOHA_MAX_RANKS];
  int32_t rc[ALOHA_MAX_RANKS][4];  // their rectangles
  RayRec recs[kFusedChunk * kFPix];  // the chunk's (rank, pixel) rays, set up
  int wcnt[kBlock / 64];
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

// Transparent black (no rank reaches the pixel) at frame pixel `pixel`
// (streaming stores, as the per-rank fill).
__device__ __forceinline__ void fused_clear(const MarchArgs &a, long long pixel) {
  if (a.rgba) store_nt(a.rgba + pixel, make_float4(0.f, 0.f, 0.f, 0.f));
  if (a.rgba8) store_nt(a.rgba8 + pixel, 0u);
  if (a.nsamples) store_nt(a.nsamples + pixel, 0u);
}

// One fused 16x4 tile at (bx, by): rank list, chunks of (rank, pixel) items
// marched by lane groups, in-order blend, output.  The TF must be staged in
// s_tf before the call; every thread of the block calls it.  Safe to call
// repeatedly on the same `fs` when a barrier separates the calls.
template <int G, bool kU8, class Src>
__device__ __forceinline__ void fused_tile(const MarchArgs &a, const Src &src, FusedStage &fs,
                                           const MarchTables &tb, TileShape ts, int bx, int by) {
  const int tw = ts.w(), th = ts.h();
  const int t = threadIdx.x;
  if (t == 0) {
    int n = 0;
    for (int k = 0; k < a.P; ++k) {
      const int s = src.order(k);
      const int32_t *rc = src.rect(s);
      if (bx < rc[2] && bx + tw > rc[0] && by < rc[3] && by + th > rc[1]) {
        for (int i = 0; i < 4; ++i) fs.rc[n][i] = rc[i];
        fs.ranks[n++] = s;
      }
    }
    fs.nranks = n;
  }
  if (t < kFPix) fs.ns[t] = 0u;
  __syncthreads();
  const int nranks = fs.nranks;  // block-uniform
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  const float inv_spp = a.inv_spp;
  const int own = ts.index(t & (tw - 1), (t >> ts.wl) & (th - 1));  // pixel of thread t < kFPix
  for (int c0 = 0; c0 < nranks; c0 += kFusedChunk) {
    const int nc = min(kFusedChunk, nranks - c0);
    __syncthreads();  // previous chunk blended
    const int items = nc * kFPix;
    // a rank whose bricks the pixel's ray misses contributes exact zeros
    for (int i = t; i < items; i += kBlock) fs.col[i / kFPix][i % kFPix] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t == 0) fs.next = 0;
    auto store = [&](const ListRay &ray, float4 acc, float, uint32_t ns, long long) {
      const int item = ray.tag, r = item / kFPix, idx = item - r * kFPix;
      fs.col[r][idx] = make_float4((acc.x * 1.f) * inv_spp, (acc.y * 1.f) * inv_spp,
                                   (acc.z * 1.f) * inv_spp, (acc.w * 1.f) * inv_spp);
      if (ns) atomicAdd(&fs.ns[idx], ns);
    };
    // every item's ray set up by its own thread; the rays that reach a brick
    // of their rank go to the list (the others contribute the zeros above)
    const int count = compact_records(items, fs.recs, fs.wcnt, [&](int item, RayState &st) {
      const int r = item / kFPix, idx = item - r * kFPix;
      const int slot = fs.ranks[c0 + r];
      int lx, ly;
      ts.pixel(idx, lx, ly);
      const int px = bx + lx, py = by + ly;
      if (px >= a.W || py >= a.H || !in_rc(fs.rc[c0 + r], px, py) || !island_pixel(a, px, py))
        return;
      begin_ray(a, tb, src.org(), src.dir00(), src.du(), src.dv(), st, ListRay{slot, px, py, item},
                true, store);
    });
    march_list<G, true, kU8>(
        a, tb, src.org(), src.dir00(), src.du(), src.dv(),
        [&](bool want) { return want ? atomicAdd(&fs.next, 1) : 0; }, count,
        [&](int k) { return fs.recs[k]; }, store);
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
  const int px = bx + (t & (tw - 1)), py = by + (t >> ts.wl);
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
template <int G, bool kU8>
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
  const int item = (int)blockIdx.x;  // row-major tiles of the box
  const MarchTables tb = load_tables<kU8>(a, s_tf);
  fused_tile<G, kU8>(a, src, fs, tb, TileShape{kFusedWl, kFusedHl}, (u.x + item % u.z) * kFW,
                (u.y + item / u.z) * kFH);
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
  bx = (bx4[0] + k % bx4[2]) << b.tile_wl;
  by = (bx4[1] + k / bx4[2]) << (6 - b.tile_wl);
}

__device__ __forceinline__ void batch_fill(const BatchArgs &b, int fb, int nfb) {
  fill_rows(
      b.m.W, b.m.H, b.V, fb, nfb,
      [&](int view, int &x0, int &y0, int &x1, int &y1) {
        const int32_t *bx4 = b.boxes + 4 * view;
        const int wl = b.tile_wl, hl = 6 - b.tile_wl;
        x0 = bx4[0] << wl, y0 = bx4[1] << hl;
        x1 = (bx4[0] + bx4[2]) << wl, y1 = (bx4[1] + bx4[3]) << hl;
      },
      [&](int view, int px, int py) {
        fused_clear(b.m, (long long)view * b.m.W * b.m.H + px + (long long)b.m.W * py);
      });
}

template <int G, bool kPersistent, bool kU8>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchArgs b) {
  extern __shared__ float4 s_tf[];
  __shared__ FusedStage fs;
  __shared__ int s_item;
  if (!kPersistent) {
    if ((int)blockIdx.x >= b.total_active) {
      batch_fill(b, (int)blockIdx.x - b.total_active, b.fill_blocks);
      return;
    }
    const int item = (int)blockIdx.x;
    const MarchTables tb = load_tables<kU8>(b.m, s_tf);
    int view, bx, by;
    batch_active(b, item, view, bx, by);
    fused_tile<G, kU8>(b.m, ViewSrc{b, view}, fs, tb, TileShape{b.tile_wl, 6 - b.tile_wl}, bx, by);
    return;
  }
  // XCD x works through part x of every view (a contiguous band of each
  // view's tiles, so its L2 holds the bricks of that band), then helps the
  // other XCDs with what is left of theirs
  const MarchTables tb = load_tables<kU8>(b.m, s_tf);
  const int x0 = (int)(blockIdx.x % kXcds);
  for (int q = 0; q < kXcds; ++q) {
    const int x = (x0 + q) % kXcds;
    const int32_t *xo = b.xoff + x * (b.V + 1);
    const int total = xo[b.V];
    for (int it = 0; it <= total; ++it) {  // bounded: each pass takes a fresh index
      if (threadIdx.x == 0) s_item = (int)atomicAdd(b.counter + x, 1u);
      __syncthreads();
      const int i = s_item;
      __syncthreads();  // every wave has read s_item; the previous tile's epilogue is done
      if (i >= total) break;
      int lo = 0, hi = b.V;  // the last view whose part starts at or before i
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xo[mid] <= i) lo = mid;
        else hi = mid;
      }
      const int n = b.box_off[lo + 1] - b.box_off[lo];
      const int item = b.box_off[lo] + (int)((long long)n * x / kXcds) + (i - xo[lo]);
      int view, bx, by;
      batch_active(b, item, view, bx, by);
      fused_tile<G, kU8>(b.m, ViewSrc{b, view}, fs, tb, TileShape{b.tile_wl, 6 - b.tile_wl}, bx, by);
    }
  }
  batch_fill(b, (int)blockIdx.x, (int)gridDim.x);
}

// Work-counter reset as a kernel node (graph-capturable without a memset node).
__global__ void reset_counter_kernel(uint32_t *c) { c[threadIdx.x] = 0u; }

// Decode + pad one brick: dst element (i,j,k), i in [-1, n+2], reads the
// clamp-to-edge source voxel.  f32 storage: u8/u16 decode as Volume::get
// (u/255.f, u/65535.f); u8 storage (u8 source only): the bytes as they are.
__global__ __launch_bounds__(256) void pad_brick_kernel(const void *src, int dtype, int nx,
                                                        int ny, int nz, void *dst_base,
                                                        int store_dtype, int px, int py,
                                                        long long total) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % px) - kPadLo;
    const int j = (int)((e / px) % py) - kPadLo;
    const int k = (int)(e / ((long long)px * py)) - kPadLo;
    const long long s = ((long long)clampi(k, 0, nz - 1) * ny + clampi(j, 0, ny - 1)) * nx +
                        clampi(i, 0, nx - 1);
    if (store_dtype == ALOHA_U8) {
      ((uint8_t *)dst_base)[e] = ((const uint8_t *)src)[s];
      continue;
    }
    float v;
    if (dtype == ALOHA_U8) v = (float)((const uint8_t *)src)[s] / 255.f;
    else if (dtype == ALOHA_U16) v = (float)((const uint16_t *)src)[s] / 65535.f;
    else v = ((const float *)src)[s];
    ((float *)dst_base)[e] = v;
  }
}

}  // namespace

// fill blocks of the per-rank and fused launches: a few per CU, each
// striding over tiles
constexpr int kFillBlocks = 1024;

// f(std::integral_constant<int, G>, std::bool_constant<kU8>) for G = lanes
// (1, 2, 4 or 8 lanes per ray) and kU8 = u8 (the u8 decode compiled in: a
// brick of the launch keeps u8 voxels)
template <class F>
hipError_t with_schedule(int lanes, bool u8, const F &f) {
  auto go = [&](auto g) { return u8 ? f(g, std::true_type{}) : f(g, std::false_type{}); };
  switch (lanes) {
    case 1: return go(std::integral_constant<int, 1>{});
    case 2: return go(std::integral_constant<int, 2>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default: return hipErrorInvalidValue;
  }
}

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
  const size_t lds = march_tables_bytes(a);
  return with_schedule(group, a.u8_lut != 0, [&](auto g, auto u8) {
    march_composite_group_kernel<g(), u8()><<<grid, kBlock, lds, s>>>(a, u, fill_blocks);
    return hipGetLastError();
  });
}

hipError_t launch_batch(const BatchArgs &b, int lanes, int persistent_blocks, hipStream_t s) {
  const size_t lds = march_tables_bytes(b.m);
  if (persistent_blocks > 0) {
    reset_counter_kernel<<<1, kXcds, 0, s>>>(b.counter);
    const dim3 grid(persistent_blocks);
    return with_schedule(lanes, b.m.u8_lut != 0, [&](auto g, auto u8) {
      batch_kernel<g(), true, u8()><<<grid, kBlock, lds, s>>>(b);
      return hipGetLastError();
    });
  }
  const dim3 grid((unsigned)(b.total_active + b.fill_blocks));
  return with_schedule(lanes, b.m.u8_lut != 0, [&](auto g, auto u8) {
    batch_kernel<g(), false, u8()><<<grid, kBlock, lds, s>>>(b);
    return hipGetLastError();
  });
}

hipError_t launch_march_footprint(const MarchArgs &a, hipStream_t s) {
  const dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileH - 1) / kTileH, a.P);
  const size_t lds = sizeof(float4) * (size_t)(a.tf_n + 4) + sizeof(float) * kU8Lut;
  march_kernel<true><<<grid, kBlock, lds, s>>>(a);
  return hipGetLastError();
}

hipError_t ranks_occupancy(int group, bool u8, size_t lds, int *blocks_per_cu) {
  return with_schedule(group, u8, [&](auto g, auto k) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                        march_persist_kernel<g(), k()>, 64, lds);
  });
}

hipError_t launch_march_ranks(const MarchArgs &a, const RanksArgs &q, int group, int waves,
                              hipStream_t s) {
  const size_t lds = ranks_lds_bytes(a);
  return with_schedule(group, a.u8_lut != 0, [&](auto g, auto u8) {
    march_persist_kernel<g(), u8()><<<waves, 64, lds, s>>>(a, q);
    return hipGetLastError();
  });
}

hipError_t launch_pad_brick(const void *src, int dtype, const int32_t n[3], void *dst_base,
                            int store_dtype, int32_t px, int32_t py, int32_t pz, hipStream_t s) {
  if (store_dtype == ALOHA_U8 && dtype != ALOHA_U8) return hipErrorInvalidValue;
  const long long total = (long long)px * py * pz;
  const long long blocks = (total + 255) / 256;
  const int grid = (int)(blocks < 65536 ? blocks : 65536);
  pad_brick_kernel<<<grid, 256, 0, s>>>(src, dtype, n[0], n[1], n[2], dst_base, store_dtype, px,
                                        py, total);
  return hipGetLastError();
}

}  // namespace aloha
