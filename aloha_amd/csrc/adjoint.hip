// adjoint.hip -- differentiable form of the per-brick ray march: a forward
// kernel and its hand-written adjoint, giving dLoss/dTF and dLoss/dDensity
// for an image loss (no reference counterpart; the forward restates
// integrateBrick, deviceCode.cpp:256-298, for one rank marching its bricks).
//
// Two deliberate differences from the plain raygen kernel make the march
// reversible:
//   * samples sit on the closed-form lattice t_k = dt*(j + (k0+k)) (the
//     lattice firstSampleT starts, evaluated per index) instead of repeated
//     t += dt, so the adjoint can visit them in reverse order exactly;
//   * the per-sample opacity is clamped to 0.9999 so transmittance can be
//     recovered backward by division.
// The adjoint stores nothing per sample: from the forward's final colour
// and alpha (and per-pixel sample count) it re-derives each brick visit,
// walks its samples from last to first, recovers T_k = T_{k+1}/(1-a_k)
// (Weiss & Westermann, "Differentiable Direct Volume Rendering", IEEE VIS
// 2021, the inversion method) and back-propagates through compositing,
// shading (N = g/|g+1e-4|), the TF lookup and all seven trilinear taps.
// dTF is summed in registers over each run of a ray's samples between the
// same two TF entries, then in LDS per workgroup, and flushed with one global
// atomic per touched entry; dDensity with global float atomics on the brick's
// unpadded voxel grid (clamped corners receive the weight of the clamped-away
// neighbour, exactly as the forward reads them), summed per voxel of the
// shared 20-voxel footprint and then per workgroup in an LDS hash before the
// atomics.  The forward marches a ray with a lane group (4 lanes, 4 lattice
// samples a step); the adjoint keeps one lane per ray (its reverse walk is a
// sequential recurrence).
#include <climits>
#include <hip/hip_runtime.h>

#include "device_types.h"
#include "march_device.h"

namespace aloha {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr int kTile = 16;

struct Visit {
  int bid;        // brick id
  int slot;       // its index in the brick list (a.brick_ids, a.dvox)
  float *grad;    // its density-gradient grid (may be null)
  float t0, t1;   // clipped interval
  float j;        // lattice jitter
  float k0;       // first lattice index
};

// The brick visit sequence of renderFrame's do/while for one ray: closest
// reported hit among the rank's bricks, ties to the lower slot.  Returns the
// number of visits (<= max_visits) and fills vis[] if non-null.
__device__ int visits(const DiffArgs &a, const float org[3], const float d[3], uint32_t pixel,
                      int max_visits, Visit *vis, int want) {
  uint32_t rng = rng_seed(pixel, 0u);
  float tmin = 0.f;
  int nv = 0;
  while (nv < max_visits) {
    int hit = -1, hslot = 0;
    float *hgrad = nullptr;
    float best = 0.f, bt0 = 0.f, bt1 = 0.f;
    for (int k = 0; k < a.nb; ++k) {
      const int bid = a.brick_ids[k];  // k is wave-uniform: scalar loads
      const DevBrick &db = a.bricks[bid];
      float t0, t1, rep;
      if (!slab(org, d, tmin, 1e10f, db.space_lo, db.space_hi, t0, t1)) continue;
      if (t0 > tmin) rep = t0;
      else if (t1 > tmin) rep = t1;
      else continue;
      if (hit < 0 || rep < best) {
        hit = bid;
        hslot = k;
        hgrad = a.dvox[k];
        best = rep;
        bt0 = t0;
        bt1 = t1;
      }
    }
    if (hit < 0) break;
    const float j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(rng);
    if (nv == want || want < 0) {
      const float dt = a.dt;
      const float segs = floorf((bt0 - dt * j) / dt);
      const float tf = dt * (j + segs);
      Visit v{hit, hslot, hgrad, bt0, bt1, j, tf < bt0 ? segs + 1.f : segs};
      if (vis) vis[want < 0 ? nv : 0] = v;
    }
    ++nv;
    tmin = bt1 + 1e-20f;
  }
  return nv;
}

__device__ __forceinline__ float lattice_t(const DiffArgs &a, const Visit &v, int k) {
  return a.dt * (v.j + (v.k0 + (float)k));
}

__device__ __forceinline__ void ray_dir(const DiffArgs &a, int view, int px, int py, float d[3]) {
  const float *c = a.cam[view];
  const float su = (float)px + .5f, sv = (float)py + .5f;
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = (c[3 + k] + su * c[6 + k]) + sv * c[9 + k];
  const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d[k] = d[k] / len;
    if (fabsf(d[k]) < 1e-5f) d[k] = 1e-5f;
  }
}

__device__ __forceinline__ BrickView view_of(const DevBrick &db) { return brick_view(db); }

struct Shaded {
  Sample s;
  float4 tf;  // TF colour, opacity clamped
  float shade;
  float len;
};

__device__ __forceinline__ Shaded shade_sample(const DiffArgs &a, const float4 *s_tf,
                                               const DevBrick &db, const float org[3],
                                               const float d[3], float t) {
  const BrickView b = view_of(db);
  const float x = (org[0] + t * d[0]) - db.vox_lo[0];
  const float y = (org[1] + t * d[1]) - db.vox_lo[1];
  const float z = (org[2] + t * d[2]) - db.vox_lo[2];
  Shaded r;
  if (!sample_fast(b, x, y, z, r.s)) r.s = sample_exact<false>(b, nullptr, db, x, y, z);
  const float ex = r.s.gx + 1e-4f, ey = r.s.gy + 1e-4f, ez = r.s.gz + 1e-4f;
  r.len = sqrtf(ex * ex + ey * ey + ez * ez);
  const float nx = r.s.gx / r.len, ny = r.s.gy / r.len, nz = r.s.gz / r.len;
  r.tf = tf_lookup(s_tf, a.tf_n, r.s.val);
  r.tf.w = fminf(r.tf.w, a.alpha_max);
  r.shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
  return r;
}

__device__ void load_tf(const DiffArgs &a, float4 *s_tf) {
  for (int i = threadIdx.x; i < a.tf_n + 4; i += kBlock)
    s_tf[i] = (i >= 2 && i < a.tf_n + 2) ? a.tf[i - 2] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// The forward with lane groups: G lanes per ray evaluate G lattice samples
// of its segment at once (t_k depends on k alone, so the lanes need no
// chain), then every lane of the group applies them in order with the
// sequential recurrence and the loop test (t < t1 && alpha < .99) before
// each: the restated march's arithmetic in its order (integrateBrick with the
// lattice t and the opacity clamp), checked bit for bit against the oracle.  Each wave owns one 8x8 tile of
// the 16x16 block; a group takes the tile's next pixel as soon as its ray is
// done.  States: kIdle (no ray), kPick (closest brick hit from tmin next),
// kMarch (stepping the segment), kOver (the tile is done).
template <int G>
__global__ __launch_bounds__(kBlock) void diff_forward_group_kernel(const DiffArgs a) {
  enum : int { kIdle = 0, kPick = 1, kMarch = 2, kOver = 3 };
  extern __shared__ float4 s_tf[];
  load_tf(a, s_tf);
  __syncthreads();
  const int view = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gl = lane % G, gbase = lane - gl;
  const int tx = blockIdx.x * kTile + (wave & 1) * 8, ty = blockIdx.y * kTile + (wave >> 1) * 8;
  const float *org = a.cam[view];
  const float dt = a.dt;
  int state = kIdle, next = 0, px = 0, py = 0;  // next: wave-uniform
  float d[3] = {1.f, 0.f, 0.f};
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t ns = 0, rng = 0;
  float tmin = 0.f, t1 = 0.f, j = 0.f, k0 = 0.f;
  int kk = 0;  // the segment's next lattice step
  BrickView bv{nullptr, 0u, 0u, 0, 0, 0};
  float lo0 = 0.f, lo1 = 0.f, lo2 = 0.f;
  while (__any(state != kOver)) {
    // a group without a ray takes the tile's next pixel
    const bool want = state == kIdle && gl == 0;
    const uint64_t m = __ballot(want);
    int k = next + __popcll(m & ((1ull << lane) - 1ull));
    next += __popcll(m);
    k = group_bcast_i<G, 0>(k, gbase);
    if (state == kIdle) {
      if (k >= 64) {
        state = kOver;
      } else {
        px = tx + (k & 7);
        py = ty + (k >> 3);
        if (px < a.W && py < a.H) {
          ray_dir(a, view, px, py, d);
          rng = rng_seed((uint32_t)(px + a.W * py), 0u);
          c = make_float4(0.f, 0.f, 0.f, 0.f);
          ns = 0;
          tmin = 0.f;
          state = kPick;
        }
      }
    }
    if (state == kPick) {  // closest reported hit from tmin, ties to the lower slot
      int hit = -1;
      float best = 0.f, bt0 = 0.f, bt1 = 0.f;
      for (int kb = 0; kb < a.nb; ++kb) {
        const DevBrick &db = a.bricks[a.brick_ids[kb]];  // kb is wave-uniform: scalar loads
        float s0, s1, rep;
        if (!slab(org, d, tmin, 1e10f, db.space_lo, db.space_hi, s0, s1)) continue;
        if (s0 > tmin) rep = s0;
        else if (s1 > tmin) rep = s1;
        else continue;
        if (hit < 0 || rep < best) {
          hit = a.brick_ids[kb];
          best = rep;
          bt0 = s0;
          bt1 = s1;
        }
      }
      if (hit < 0) {
        if (gl == 0) {
          const size_t out = (size_t)view * a.W * a.H + px + (size_t)a.W * py;
          a.rgba[out] = c;
          a.nsamples[out] = ns;
        }
        state = kIdle;
      } else {
        j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(rng);
        const float segs = floorf((bt0 - dt * j) / dt);
        const float tfirst = dt * (j + segs);
        k0 = tfirst < bt0 ? segs + 1.f : segs;
        t1 = bt1;
        kk = 0;
        const DevBrick &db = a.bricks[hit];
        bv = brick_view(db);
        lo0 = db.vox_lo[0];
        lo1 = db.vox_lo[1];
        lo2 = db.vox_lo[2];
        tmin = bt1 + 1e-20f;
        state = kMarch;
      }
    }
    // kMarch: lattice samples kk .. kk + G - 1 of the segment
    const bool go = state == kMarch;
    const float t = dt * (j + (k0 + (float)(kk + gl)));
    float4 tf = make_float4(0.f, 0.f, 0.f, 0.f);
    float shade = -1.f;  // < 0: no sample (an evaluated shade is >= 0.1 or NaN)
    if (go && t < t1) {
      const float x = (org[0] + t * d[0]) - lo0;
      const float y = (org[1] + t * d[1]) - lo1;
      const float z = (org[2] + t * d[2]) - lo2;
      Sample sm;
      if (!sample_fast(bv, x, y, z, sm)) sm = sample_exact<false>(bv, nullptr, a.bricks[0], x, y, z);
      const float ex = sm.gx + 1e-4f, ey = sm.gy + 1e-4f, ez = sm.gz + 1e-4f;
      const float len = sqrtf(ex * ex + ey * ey + ez * ez);
      const float nx = sm.gx / len, ny = sm.gy / len, nz = sm.gz / len;
      tf = tf_lookup(s_tf, a.tf_n, sm.val);
      tf.w = fminf(tf.w, a.alpha_max);
      shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
    }
    bool live = go;
    static_for<0, G>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      const float tx_ = group_bcast<G, J>(tf.x, gbase), ty_ = group_bcast<G, J>(tf.y, gbase);
      const float tz_ = group_bcast<G, J>(tf.z, gbase), tw_ = group_bcast<G, J>(tf.w, gbase);
      const float sh = group_bcast<G, J>(shade, gbase);
      live = live && !(sh < 0.f) && c.w < .99f;
      if (live) {
        ++ns;
        const float w = (1.f - c.w) * tw_;
        c.x = c.x + (w * tx_) * sh;
        c.y = c.y + (w * ty_) * sh;
        c.z = c.z + (w * tz_) * sh;
        c.w = c.w + w;
      }
    });
    if (go) {
      kk += G;
      const float tn = dt * (j + (k0 + (float)kk));
      if (!(tn < t1 && c.w < .99f)) {  // segment over
        if (c.w < .99f) {
          state = kPick;
        } else {
          if (gl == 0) {
            const size_t out = (size_t)view * a.W * a.H + px + (size_t)a.W * py;
            a.rgba[out] = c;
            a.nsamples[out] = ns;
          }
          state = kIdle;
        }
      }
    }
  }
}

// Where a brick visit's density gradient goes.  Global float atomics run at
// about 20 G/s on MI355X whatever the address pattern, and neighbouring rays
// and samples hit the same voxels many times, so each workgroup first sums
// into an LDS hash table of (brick slot, voxel) -> partial gradient and, at
// its end, issues one global atomic per distinct voxel.  A brick whose voxel
// index does not fit the 24-bit key, and any update whose short probe run is
// full, goes straight to a global atomic.  Sums are reassociated either way
// (float atomics are unordered).  Sized by measurement on the 256^3 x 8
// views workload: 8192 slots (64 KB, two workgroups per CU) with runs of two
// probes beat 4096 slots with eight (fewer global fall-throughs, shorter
// chains of dependent LDS compare-and-swaps): 5.2 vs 6.7 ms forward+adjoint.
constexpr int kCacheBits = 13;
constexpr int kCacheSlots = 1 << kCacheBits;
constexpr int kCacheProbes = 2;
constexpr uint32_t kCacheEmpty = 0xFFFFFFFFu;

struct GradSink {
  float *grad;     // the brick's unpadded gradient grid
  uint32_t *keys;  // LDS table keys (null: no caching for this brick)
  float *vals;     // LDS table partial sums
  uint32_t tag;    // brick slot << 24
};

// LDS and global views of the table and the grid, so each add is an LDS or a
// global atomic (a generic pointer would make both flat atomics)
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(1))) float glb_float;

__device__ __forceinline__ void sink_add(const GradSink &s, uint32_t idx, float v) {
  if (s.keys) {
    const uint32_t key = s.tag | idx;
    uint32_t h = (key * 2654435761u) >> (32 - kCacheBits);
    for (int p = 0; p < kCacheProbes; ++p) {
      const uint32_t k = atomicCAS(&s.keys[h], kCacheEmpty, key);
      if (k == kCacheEmpty || k == key) {
        __hip_atomic_fetch_add((lds_float *)s.vals + h, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
      }
      h = (h + 1u) & (uint32_t)(kCacheSlots - 1);
    }
  }
  __hip_atomic_fetch_add((glb_float *)s.grad + idx, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Scatter d(loss)/d(tap value) onto the 8 corners of one trilinear tap.
__device__ __forceinline__ void scatter_tap(const GradSink &sink, const DevBrick &db, float x,
                                            float y, float z, float g) {
  if (g == 0.f) return;
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const float wx = x - fx, wy = y - fy, wz = z - fz;
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const int nx = db.n[0], ny = db.n[1], nz = db.n[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int zz = clampi(iz + k, 0, nz - 1);
    const float w_z = k ? wz : 1.f - wz;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int yy = clampi(iy + j, 0, ny - 1);
      const float w_yz = (j ? wy : 1.f - wy) * w_z;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int xx = clampi(ix + i, 0, nx - 1);
        const float w = (i ? wx : 1.f - wx) * w_yz;
        sink_add(sink, (uint32_t)(((size_t)zz * ny + yy) * nx + xx), g * w);
      }
    }
  }
}

// x index ax + scx + ii of a footprint row, without dynamic register indexing
template <int II>
__device__ __forceinline__ void add_x(float row[3], int scx, float v) {
  if (II == 0) {
    row[0] += scx ? 0.f : v;
    row[1] += scx ? v : 0.f;
  } else {
    row[1] += scx ? 0.f : v;
    row[2] += scx ? v : 0.f;
  }
}

__device__ __forceinline__ void atomic_voxel(const GradSink &sink, const DevBrick &db, int x, int y,
                                             int z, float v) {
  if (v == 0.f) return;
  x = clampi(x, 0, db.n[0] - 1);
  y = clampi(y, 0, db.n[1] - 1);
  z = clampi(z, 0, db.n[2] - 1);
  sink_add(sink, (uint32_t)(((size_t)z * db.n[1] + y) * db.n[0] + x), v);
}

// Scatter of all seven taps of one sample through the shared 20-voxel
// footprint (sample_fast's frame: R = 3 x-columns of the value tap's y/z
// pair, one extra y row, one extra z plane).  Each tap's eight corner
// weights are summed per footprint voxel in registers, so the sample issues
// at most 20 atomics instead of 56; the sum equals the tap-by-tap scatter up
// to float reassociation.  gV: d/d(value tap); gd: d/d(gradient component),
// i.e. +gd for the +0.5 tap and -gd for the -0.5 tap of each axis.  Returns
// false, scattering nothing, when the footprint does not fit.
__device__ __forceinline__ bool scatter_fast(const GradSink &sink, const DevBrick &db, float x,
                                             float y, float z, float gV, const float gd[3]) {
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
                  (bx == ax + 1) & (by == ay + 1) & (bz == az + 1) & (ax >= -1) &
                  (ax <= db.n[0]) & (ay >= -1) & (ay <= db.n[1]) & (az >= -1) & (az <= db.n[2]);
  if (!ok) return false;
  const float wxm = xm - fxm, wxc = x - fxc, wxp = xp - fxp;
  const float wym = ym - fym, wyc = y - fyc, wyp = yp - fyp;
  const float wzm = zm - fzm, wzc = z - fzc, wzp = zp - fzp;
  const float gXp = gd[0], gXm = -gd[0], gYp = gd[1], gYm = -gd[1], gZp = gd[2], gZm = -gd[2];
  float A[2][2][3];  // A[kk][jj][i]: voxel (ax + i, cy + jj, cz + kk)
  float YN[2][2];    // YN[kk][ii]: voxel (cx + ii, yn, cz + kk)
  float ZN[2][2];    // ZN[jj][ii]: voxel (cx + ii, cy + jj, zn)
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      // value tap and the two x taps share the y rows cy, cy+1 and z planes cz, cz+1
      const float wyz = (jj ? wyc : 1.f - wyc) * (kk ? wzc : 1.f - wzc);
      const float v0 = gV * wyz * (1.f - wxc), v1 = gV * wyz * wxc;
      A[kk][jj][0] = (scx ? 0.f : v0) + gXm * wyz * (1.f - wxm);
      A[kk][jj][1] = (scx ? v0 : v1) + gXm * wyz * wxm + gXp * wyz * (1.f - wxp);
      A[kk][jj][2] = (scx ? v1 : 0.f) + gXp * wyz * wxp;
      YN[kk][jj] = 0.f;
      ZN[kk][jj] = 0.f;
    }
  // y taps: x pair (cx, cx+1), z planes (cz, cz+1); rows ay, ay+1, ay+2
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const float wz = kk ? wzc : 1.f - wzc;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
      const float wxz = (ii ? wxc : 1.f - wxc) * wz;
      const float r0 = gYm * wxz * (1.f - wym);
      const float r1 = gYm * wxz * wym + gYp * wxz * (1.f - wyp);
      const float r2 = gYp * wxz * wyp;
      YN[kk][ii] += scy ? r0 : r2;  // the row outside the value tap's pair
      if (ii == 0) {
        add_x<0>(A[kk][0], scx, scy ? r1 : r0);
        add_x<0>(A[kk][1], scx, scy ? r2 : r1);
      } else {
        add_x<1>(A[kk][0], scx, scy ? r1 : r0);
        add_x<1>(A[kk][1], scx, scy ? r2 : r1);
      }
    }
  }
  // z taps: x pair (cx, cx+1), y rows (cy, cy+1); planes az, az+1, az+2
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) {
    const float wy = jj ? wyc : 1.f - wyc;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
      const float wxy = (ii ? wxc : 1.f - wxc) * wy;
      const float p0 = gZm * wxy * (1.f - wzm);
      const float p1 = gZm * wxy * wzm + gZp * wxy * (1.f - wzp);
      const float p2 = gZp * wxy * wzp;
      ZN[jj][ii] += scz ? p0 : p2;
      if (ii == 0) {
        add_x<0>(A[0][jj], scx, scz ? p1 : p0);
        add_x<0>(A[1][jj], scx, scz ? p2 : p1);
      } else {
        add_x<1>(A[0][jj], scx, scz ? p1 : p0);
        add_x<1>(A[1][jj], scx, scz ? p2 : p1);
      }
    }
  }
  const int yn = scy ? ay : ay + 2;
  const int zn = scz ? az : az + 2;
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int i = 0; i < 3; ++i) atomic_voxel(sink, db, ax + i, cy + jj, cz + kk, A[kk][jj][i]);
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) atomic_voxel(sink, db, cx + ii, yn, cz + kk, YN[kk][ii]);
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) atomic_voxel(sink, db, cx + ii, cy + jj, zn, ZN[jj][ii]);
  return true;
}

// Adds a run's TF gradient (entries pi and pi + 1; pi == INT_MIN: none) to
// the workgroup's LDS sums.  Consecutive samples of a ray mostly fall
// between the same two TF entries, so the backward sums them in registers
// and adds once per run: 8 LDS float atomics per run instead of per sample
// (measured on 256^3, 8 views of 512^2: fwd+adjoint 5.09 -> 4.92 ms, the
// same as with no TF atomics at all).
__device__ __forceinline__ void tf_flush(float *s_dtf, int n, int pi, const float p0[4],
                                         const float p1[4]) {
  if (pi == INT_MIN) return;
#pragma unroll
  for (int ch = 0; ch < 4; ++ch) {
    if (pi >= 0 && pi < n) atomicAdd(&s_dtf[4 * pi + ch], p0[ch]);
    if (pi + 1 >= 0 && pi + 1 < n) atomicAdd(&s_dtf[4 * (pi + 1) + ch], p1[ch]);
  }
}

__global__ __launch_bounds__(kBlock) void diff_backward_kernel(const DiffArgs a) {
  extern __shared__ float4 s_mem[];
  float4 *s_tf = s_mem;
  float *s_dtf = (float *)(s_mem + a.tf_n + 4);  // tf_n x 4
  uint32_t *s_keys = (uint32_t *)(s_dtf + 4 * a.tf_n);
  float *s_vals = (float *)(s_keys + kCacheSlots);
  load_tf(a, s_tf);
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += kBlock) s_dtf[i] = 0.f;
  for (int i = threadIdx.x; i < kCacheSlots; i += kBlock) {
    s_keys[i] = kCacheEmpty;
    s_vals[i] = 0.f;
  }
  __syncthreads();

  const int view = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = blockIdx.x * kTile + (wave & 1) * 8 + (lane & 7);
  const int py = blockIdx.y * kTile + (wave >> 1) * 8 + (lane >> 3);
  const bool live = px < a.W && py < a.H;
  uint32_t N = 0;
  size_t out = 0;
  // TF gradient of the current run of samples on one entry pair (pi, pi+1)
  int pi = INT_MIN;
  float p0[4] = {0.f, 0.f, 0.f, 0.f}, p1[4] = {0.f, 0.f, 0.f, 0.f};
  if (live) {
    out = (size_t)view * a.W * a.H + (px + a.W * py);
    N = a.nsamples[out];
  }
  if (N > 0) {
    const int pixel = px + a.W * py;
    const float *org = a.cam[view];
    float d[3];
    ray_dir(a, view, px, py, d);
    const float4 C = a.rgba[out];
    const float4 G = a.dldc[out];
    // visits and their sample counts: all but the last run to t1
    const int nv = visits(a, org, d, (uint32_t)pixel, kMaxVisits, nullptr, -1);
    int last = -1;
    uint32_t before_last = 0, acc = 0;
    for (int vi = 0; vi < nv && last < 0; ++vi) {
      Visit v;
      visits(a, org, d, (uint32_t)pixel, vi + 1, &v, vi);
      uint32_t cnt = 0;
      while (lattice_t(a, v, (int)cnt) < v.t1) ++cnt;
      if (acc + cnt >= N) {
        last = vi;
        before_last = acc;
      }
      acc += cnt;
    }
    float T = 1.f - C.w;  // transmittance after the last sample
    const float TN = T;
    float Sr = 0.f, Sg = 0.f, Sb = 0.f;  // suffix sums of later contributions
    for (int vi = last; vi >= 0; --vi) {
      Visit v;
      visits(a, org, d, (uint32_t)pixel, vi + 1, &v, vi);
      uint32_t cnt;
      if (vi == last) {
        cnt = N - before_last;
      } else {
        cnt = 0;
        while (lattice_t(a, v, (int)cnt) < v.t1) ++cnt;
      }
      const DevBrick db = a.bricks[v.bid];
      float *grad = v.grad;
      const bool keyed = (int64_t)db.n[0] * db.n[1] * db.n[2] <= (1 << 24);
      const GradSink sink{grad, keyed ? s_keys : nullptr, s_vals, (uint32_t)v.slot << 24};
      for (int k = (int)cnt - 1; k >= 0; --k) {
        const float t = lattice_t(a, v, k);
        const Shaded r = shade_sample(a, s_tf, db, org, d, t);
        const float am = r.tf.w;
        const float one_m = 1.f - am;
        const float Tk = T / one_m;
        const float cs_r = r.tf.x * r.shade, cs_g = r.tf.y * r.shade, cs_b = r.tf.z * r.shade;
        const float Gc = G.x * r.tf.x + G.y * r.tf.y + G.z * r.tf.z;
        // d/d(colour), d/d(shade), d/d(opacity)
        const float Ta = Tk * am;
        const float dcr = G.x * Ta * r.shade, dcg = G.y * Ta * r.shade, dcb = G.z * Ta * r.shade;
        const float dshade = Ta * Gc;
        const float dalpha = Tk * (G.x * cs_r + G.y * cs_g + G.z * cs_b) -
                             ((G.x * Sr + G.y * Sg + G.z * Sb) - G.w * TN) / one_m;
        Sr += Ta * cs_r;
        Sg += Ta * cs_g;
        Sb += Ta * cs_b;
        T = Tk;
        // TF entries (border entries are constants)
        const int n = a.tf_n;
        const float xt = r.s.val * (float)n - 0.5f;
        const float fi = floorf(xt);
        const float w = xt - fi;
        const int i0 = (int)fminf(fmaxf(fi, -2.f), (float)n);
        const float4 raw = tf_lookup(s_tf, n, r.s.val);
        const float da = raw.w < a.alpha_max ? dalpha : 0.f;
        const float g4[4] = {dcr, dcg, dcb, da};
        if (i0 != pi) {
          tf_flush(s_dtf, n, pi, p0, p1);
          pi = i0;
#pragma unroll
          for (int ch = 0; ch < 4; ++ch) p0[ch] = p1[ch] = 0.f;
        }
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
          p0[ch] += (1.f - w) * g4[ch];
          p1[ch] += w * g4[ch];
        }
        if (grad) {
          // value tap through the TF slope
          const float4 e0 = s_tf[i0 + 2], e1 = s_tf[i0 + 3];
          const float dv = (dcr * (e1.x - e0.x) + dcg * (e1.y - e0.y) + dcb * (e1.z - e0.z) +
                            da * (e1.w - e0.w)) * (float)n;
          // shading: s = .1 + .9|N.d|, N = g/L, L = |g + 1e-4|
          const float nx = r.s.gx / r.len, ny = r.s.gy / r.len, nz = r.s.gz / r.len;
          const float nd = nx * d[0] + ny * d[1] + nz * d[2];
          const float sg = nd > 0.f ? .9f : (nd < 0.f ? -.9f : 0.f);
          const float dN[3] = {dshade * sg * d[0], dshade * sg * d[1], dshade * sg * d[2]};
          const float gv[3] = {r.s.gx, r.s.gy, r.s.gz};
          const float dNg = dN[0] * gv[0] + dN[1] * gv[1] + dN[2] * gv[2];
          const float L3 = r.len * r.len * r.len;
          float dg[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) dg[q] = dN[q] / r.len - dNg * (gv[q] + 1e-4f) / L3;
          const float x = (org[0] + t * d[0]) - db.vox_lo[0];
          const float y = (org[1] + t * d[1]) - db.vox_lo[1];
          const float z = (org[2] + t * d[2]) - db.vox_lo[2];
          if (dv == 0.f && dg[0] == 0.f && dg[1] == 0.f && dg[2] == 0.f) continue;
          if (scatter_fast(sink, db, x, y, z, dv, dg)) continue;
          scatter_tap(sink, db, x, y, z, dv);
          scatter_tap(sink, db, x + .5f, y, z, dg[0]);
          scatter_tap(sink, db, x - .5f, y, z, -dg[0]);
          scatter_tap(sink, db, x, y + .5f, z, dg[1]);
          scatter_tap(sink, db, x, y - .5f, z, -dg[1]);
          scatter_tap(sink, db, x, y, z + .5f, dg[2]);
          scatter_tap(sink, db, x, y, z - .5f, -dg[2]);
        }
      }
    }
  }
  tf_flush(s_dtf, a.tf_n, pi, p0, p1);
  __syncthreads();
  if (a.dtf)
    for (int i = threadIdx.x; i < 4 * a.tf_n; i += kBlock)
      if (s_dtf[i] != 0.f) atomicAdd(&a.dtf[i], s_dtf[i]);
  // one global atomic per distinct cached voxel
  for (int i = threadIdx.x; i < kCacheSlots; i += kBlock) {
    const uint32_t key = s_keys[i];
    const float v = s_vals[i];
    if (key != kCacheEmpty && v != 0.f) atomicAdd(&a.dvox[key >> 24][key & 0xFFFFFFu], v);
  }
}

}  // namespace

hipError_t launch_diff(const DiffArgs &a, bool backward, hipStream_t s) {
  const dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, a.V);
  const size_t tf_lds = sizeof(float4) * (size_t)(a.tf_n + 4);
  if (backward)
    diff_backward_kernel<<<grid, kBlock,
                           tf_lds + sizeof(float) * 4 * (size_t)a.tf_n +
                               (sizeof(uint32_t) + sizeof(float)) * (size_t)kCacheSlots,
                           s>>>(a);
  else
    diff_forward_group_kernel<4><<<grid, kBlock, tf_lds, s>>>(a);
  return hipGetLastError();
}

}  // namespace aloha
