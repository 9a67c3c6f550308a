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
// The forward stores no per-sample state: from its final colour and alpha
// (and per-pixel sample count) the adjoint re-derives each brick visit,
// walks its samples from last to first, recovers T_k = T_{k+1}/(1-a_k)
// (Weiss & Westermann, "Differentiable Direct Volume Rendering", IEEE VIS
// 2021, the inversion method) and back-propagates through compositing,
// shading (N = g/|g+1e-4|), the TF lookup and all seven trilinear taps.
// The adjoint runs in stages:
//   * diff_wave_slots_kernel gives each wave of pixels (an 8x8 tile) its
//     sample slots and lists the waves with samples: the next two stages
//     launch over those only (on 256^3 x 8 views of 512^2, 13 % of the
//     tiles; the other workgroups held up the dispatch of the working ones
//     for most of the kernels' time);
//   * diff_resample_kernel evaluates every sample the forward took once more
//     (value tap and gradient, 16 bytes per slot; lane groups of 8);
//   * diff_backward_kernel, one lane per ray, walks the samples in reverse
//     and writes one record per sample (position, d/d value tap, d/d gradient,
//     value, d/d TF colour and opacity), marking each run of samples that
//     falls into one 16^3 spatial bin at its head slot;
//   * diff_run_rank_kernel / diff_scan_kernel / diff_run_list_kernel place
//     the runs into per-bin slot lists (integer atomics, one per distinct bin
//     of a wave);
//   * diff_bin_kernel scatters each bin's records onto an 18^3 tile and the
//     TF entries as 64-bit fixed-point LDS sums (clamped corners receive the
//     weight of the clamped-away neighbour, exactly as the forward reads
//     them); diff_fold_kernel folds the tiles into dDensity and the TF
//     kernels reduce the bins' partials, both in a fixed order.
// No float atomics anywhere, so the gradients are bit-reproducible.
#include <algorithm>
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
// number of visits found (<= max_visits) and sets *vis to visit `want`.
__device__ __forceinline__ int visits(const DiffArgs &a, const float org[3], const float d[3], uint32_t pixel,
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
    if (nv == want && vis) {
      const float dt = a.dt;
      const float segs = floorf((bt0 - dt * j) / dt);
      const float tf = dt * (j + segs);
      *vis = Visit{hit, hslot, hgrad, bt0, bt1, j, tf < bt0 ? segs + 1.f : segs};
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

struct Shaded {
  Sample s;
  float4 tf;  // TF colour, opacity clamped
  float raw_a;  // TF opacity before the clamp
  float shade;
  float len;
  float nd;  // N . d as the shade evaluates it (its sign: the side of the |N.d| kink)
};

// Normal, TF colour and shade of a sample (value tap and gradient), in the
// forward's arithmetic order.
__device__ __forceinline__ Shaded shade_of(const DiffArgs &a, const float4 *s_tf, const Sample &s,
                                           const float d[3]) {
  Shaded r;
  r.s = s;
  const float ex = r.s.gx + 1e-4f, ey = r.s.gy + 1e-4f, ez = r.s.gz + 1e-4f;
  r.len = sqrtf(ex * ex + ey * ey + ez * ez);
  const float nx = r.s.gx / r.len, ny = r.s.gy / r.len, nz = r.s.gz / r.len;
  r.tf = tf_lookup(s_tf, a.tf_n, r.s.val);
  r.raw_a = r.tf.w;
  r.tf.w = fminf(r.tf.w, a.alpha_max);
  r.nd = nx * d[0] + ny * d[1] + nz * d[2];
  r.shade = .1f + .9f * fabsf(r.nd);
  return r;
}

__device__ void load_tf(const DiffArgs &a, float4 *s_tf) {
  for (int i = threadIdx.x; i < a.tf_n + 4; i += kBlock)
    s_tf[i] = (i >= 2 && i < a.tf_n + 2) ? a.tf[i - 2] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// The bin of a sample at brick-texel position (x, y, z) of a brick of n0 x
// n1 x n2 voxels whose bins start at `base`, nbx x nby of them per slice.
__device__ __forceinline__ int bin_of(int base, int nbx, int nby, int n0, int n1, int n2, float x,
                                      float y, float z) {
  const int ax = clampi((int)floorf(x - .5f), -1, n0);
  const int ay = clampi((int)floorf(y - .5f), -1, n1);
  const int az = clampi((int)floorf(z - .5f), -1, n2);
  const int bx = (ax + 1) / kBinEdge, by = (ay + 1) / kBinEdge, bz = (az + 1) / kBinEdge;
  return base + (bz * nby + by) * nbx + bx;
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
// The miss result (transparent, no samples) of every pixel outside the
// views' boxes: fill block fb of nfb takes rows fb, fb + nfb, ... of the
// V x H rows.
__device__ __forceinline__ void diff_fill(const DiffArgs &a, int fb, int nfb) {
  for (int r = fb; r < a.V * a.H; r += nfb) {
    const int v = r / a.H, y = r - v * a.H;
    const int x0 = a.fbox[v][0] * kTile, x1 = (a.fbox[v][0] + a.fbox[v][2]) * kTile;
    const int y0 = a.fbox[v][1] * kTile, y1 = (a.fbox[v][1] + a.fbox[v][3]) * kTile;
    const bool band = y >= y0 && y < y1;
    for (int x = threadIdx.x; x < a.W; x += blockDim.x) {
      if (band && x >= x0 && x < x1) continue;  // a box pixel: marched
      const size_t o = (size_t)v * a.W * a.H + x + (size_t)a.W * y;
      a.rgba[o] = make_float4(0.f, 0.f, 0.f, 0.f);
      a.nsamples[o] = 0u;
    }
  }
}

// Grid: the 16x16 blocks of the views' boxes (a.block_off), then blocks
// that fill in the miss result outside them -- launching only the blocks
// that can reach a brick (on 256^3 seen from 2.5 diagonals, a quarter of a
// view's blocks), so that empty blocks do not hold up the dispatch of the
// others.
template <int G>
__global__ __launch_bounds__(kBlock) void diff_forward_group_kernel(const DiffArgs a) {
  enum : int { kIdle = 0, kPick = 1, kMarch = 2, kOver = 3 };
  const int nbox = a.block_off[a.V];
  if ((int)blockIdx.x >= nbox) {  // block-uniform
    diff_fill(a, (int)blockIdx.x - nbox, (int)gridDim.x - nbox);
    return;
  }
  extern __shared__ float4 s_tf[];
  load_tf(a, s_tf);
  __syncthreads();
  int view = 0;
  while ((int)blockIdx.x >= a.block_off[view + 1]) ++view;
  const int bl = (int)blockIdx.x - a.block_off[view], bnx = a.fbox[view][2];
  const int bx = a.fbox[view][0] + bl % bnx, by = a.fbox[view][1] + bl / bnx;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gl = lane % G, gbase = lane - gl;
  const int tx = bx * kTile + (wave & 1) * 8, ty = by * kTile + (wave >> 1) * 8;
  const float *org = a.cam[view];
  const float dt = a.dt;
  int state = kIdle, next = 0, px = 0, py = 0;  // next: wave-uniform
  float d[3] = {1.f, 0.f, 0.f};
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t ns = 0, rng = 0;
  float tmin = 0.f, t1 = 0.f, j = 0.f, k0 = 0.f;
  int kk = 0;  // the segment's next lattice step
  BrickView bv{nullptr, 0u, 0u, 0, 0, 0, nullptr, false};
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
        bv = brick_view(db, nullptr);  // f32 bricks only (checked by the host)
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
      if (!sample_fast<false>(bv, x, y, z, sm))
        sm = sample_exact<false, false>(bv, nullptr, a.bricks[0], x, y, z);
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

// Where a density-gradient contribution goes: a sink takes voxel (x, y, z)
// of the brick's unpadded grid (already clamped to it) and a value.
//
// FixedTile: the bin's 18^3 LDS tile of 64-bit fixed-point sums.  Integer
// addition is associative, so the sums -- and the gradient -- do not depend
// on the order the lanes add in: the adjoint is bit-reproducible.  (LDS
// float atomics are also several times slower than integer ones on this
// GPU.)  Each value is rounded to a multiple of 2^-scale once; the bin's
// scale keeps the largest possible sum below 2^61.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

__device__ __forceinline__ long long to_fixed(float v, double scale) {
  return __double2ll_rn((double)v * scale);
}

struct FixedTile {
  unsigned long long *tile;  // LDS, kTileVox
  int x0, y0, z0;            // the tile's first voxel per axis
  double scale;              // 2^S of the bin
  __device__ __forceinline__ void add(int x, int y, int z, float v) const {
    const int lx = x - x0, ly = y - y0, lz = z - z0;
    // every voxel a bin's sample reads lies in its tile (the bin is keyed
    // by the sample's lowest corner); anything else is dropped, not written
    // out of bounds
    if ((unsigned)lx < (unsigned)kTileEdge && (unsigned)ly < (unsigned)kTileEdge &&
        (unsigned)lz < (unsigned)kTileEdge)
      add_at((lz * kTileEdge + ly) * kTileEdge + lx, v);
  }
  // tile voxel i (in range: the caller checked it)
  __device__ __forceinline__ void add_at(int i, float v) const {
    __hip_atomic_fetch_add((lds_u64 *)tile + i, (unsigned long long)to_fixed(v, scale),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

// Scatter d(loss)/d(tap value) onto the 8 corners of one trilinear tap.
template <class Sink>
__device__ __forceinline__ void scatter_tap(const Sink &sink, const DevBrick &db, float x, float y,
                                            float z, float g) {
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
        sink.add(xx, yy, zz, g * w);
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

// Scatter of all seven taps of one sample through the shared 20-voxel
// footprint (sample_fast's frame: R = 3 x-columns of the value tap's y/z
// pair, one extra y row, one extra z plane).  Each tap's eight corner
// weights are summed per footprint voxel in registers, so the sample issues
// at most 20 atomics instead of 56; the sum equals the tap-by-tap scatter up
// to float reassociation.  gV: d/d(value tap); gd: d/d(gradient component),
// i.e. +gd for the +0.5 tap and -gd for the -0.5 tap of each axis.  Returns
// false, scattering nothing, when the footprint does not fit.
template <class Sink>
__device__ __forceinline__ bool scatter_fast(const Sink &sink, const DevBrick &db, float x, float y,
                                             float z, float gV, const float gd[3]) {
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
  // The footprint's clamped coordinates as tile offsets: x columns ax ..
  // ax + 2 (cx, cx + 1 among them), y rows cy, cy + 1 and yn (the row
  // outside the value tap's pair), z planes cz, cz + 1 and zn.  Every one
  // lies in the bin's tile (the bin is keyed by the sample's lowest corner);
  // checked once here, and a footprint that left it would take the per-tap
  // scatter, which drops what falls outside.
  const int yn = scy ? ay : ay + 2;
  const int zn = scz ? az : az + 2;
  const int n0 = db.n[0] - 1, n1 = db.n[1] - 1, n2 = db.n[2] - 1;
  const int ox0 = clampi(ax, 0, n0) - sink.x0, ox1 = clampi(ax + 1, 0, n0) - sink.x0,
            ox2 = clampi(ax + 2, 0, n0) - sink.x0;
  const int oy0 = clampi(cy, 0, n1) - sink.y0, oy1 = clampi(cy + 1, 0, n1) - sink.y0,
            oyn = clampi(yn, 0, n1) - sink.y0;
  const int oz0 = clampi(cz, 0, n2) - sink.z0, oz1 = clampi(cz + 1, 0, n2) - sink.z0,
            ozn = clampi(zn, 0, n2) - sink.z0;
  constexpr unsigned kE = (unsigned)kTileEdge;
  if (!(((unsigned)ox0 < kE) & ((unsigned)ox1 < kE) & ((unsigned)ox2 < kE) &
        ((unsigned)oy0 < kE) & ((unsigned)oy1 < kE) & ((unsigned)oyn < kE) &
        ((unsigned)oz0 < kE) & ((unsigned)oz1 < kE) & ((unsigned)ozn < kE)))
    return false;
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
  // tile rows (z plane, y row) of the footprint, then the x columns
  const int ozs[2] = {oz0 * kTileEdge, oz1 * kTileEdge};
  const int oys[2] = {oy0, oy1};
  const int ox[3] = {ox0, ox1, ox2};
  const int oxc[2] = {scx ? ox1 : ox0, scx ? ox2 : ox1};  // columns cx, cx + 1
  auto put = [&](int i, float v) {
    if (v != 0.f) sink.add_at(i, v);
  };
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int row = (ozs[kk] + oys[jj]) * kTileEdge;
#pragma unroll
      for (int i = 0; i < 3; ++i) put(row + ox[i], A[kk][jj][i]);
    }
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const int row = (ozs[kk] + oyn) * kTileEdge;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) put(row + oxc[ii], YN[kk][ii]);
  }
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) {
    const int row = (ozn * kTileEdge + oys[jj]) * kTileEdge;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) put(row + oxc[ii], ZN[jj][ii]);
  }
  return true;
}

// Samples of visit v on the lattice before its exit t1, at most limit.
__device__ __forceinline__ uint32_t visit_count(const DiffArgs &a, const Visit &v, uint32_t limit) {
  uint32_t cnt = 0;
  while (cnt < limit && lattice_t(a, v, (int)cnt) < v.t1) ++cnt;
  return cnt;
}

// Slots of the waves: each wave of the 16x16-pixel blocks (an 8x8 tile of
// one view) gets 64 x its longest sample count.  A block takes its four
// waves' slots with one atomic on one of kSlotCounters counters (block id
// mod kSlotCounters, so no counter serves more than a few hundred blocks);
// counter c's slots start where the host puts its region once it has read
// the counters.  wave_off[w] = c << 56 | the wave's offset in the region.
// Which slots hold which samples does not change any sum.
__device__ __forceinline__ int wave_id(const DiffArgs &a, int view, int bx, int by, int wave) {
  const int gx = (a.W + kTile - 1) / kTile, gy = (a.H + kTile - 1) / kTile;
  return ((view * gy + by) * gx + bx) * 4 + wave;
}
// Inverse of wave_id: the view, 16x16 block and wave of wave w.
__device__ __forceinline__ void wave_of(const DiffArgs &a, int w, int &view, int &bx, int &by,
                                        int &wave) {
  const int gx = (a.W + kTile - 1) / kTile, gy = (a.H + kTile - 1) / kTile;
  wave = w & 3;
  int t = w >> 2;
  bx = t % gx;
  t /= gx;
  by = t % gy;
  view = t / gy;
}
// The first slot of pixel (px, py)'s ray (its place in its wave's tile added).
__device__ __forceinline__ unsigned long long ray_slot(const DiffArgs &a, const DiffBins &b,
                                                      int view, int px, int py) {
  const int w = wave_id(a, view, px / kTile, py / kTile, ((px >> 3) & 1) + 2 * ((py >> 3) & 1));
  const unsigned long long wo = b.wave_off[w];
  return b.region[wo >> 56] + (wo & ((1ull << 56) - 1ull)) + (unsigned)((px & 7) + 8 * (py & 7));
}

__global__ __launch_bounds__(kBlock) void diff_wave_slots_kernel(const DiffArgs a, const DiffBins b) {
  __shared__ unsigned long long s_size[4];
  const int view = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = blockIdx.x * kTile + (wave & 1) * 8 + (lane & 7);
  const int py = blockIdx.y * kTile + (wave >> 1) * 8 + (lane >> 3);
  uint32_t n = 0;
  if (px < a.W && py < a.H) n = a.nsamples[(size_t)view * a.W * a.H + px + (size_t)a.W * py];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) n = max(n, (uint32_t)__shfl_xor((int)n, o));
  if (lane == 0) s_size[wave] = 64ull * n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long total = s_size[0] + s_size[1] + s_size[2] + s_size[3];
    const int c = (int)((blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) %
                        kSlotCounters);
    // the walks launch over the waves with samples only: most tiles of a
    // view miss the bricks, and workgroups that find nothing to do still
    // hold up the dispatch of the ones behind them.  The same atomic takes
    // the block's slots and its places in counter c's list of such waves.
    const unsigned long long nl = (s_size[0] > 0) + (s_size[1] > 0) + (s_size[2] > 0) +
                                  (s_size[3] > 0);
    const unsigned long long old =
        total ? atomicAdd(&b.slot_ctr[c], total | nl << kLiveShift) : 0ull;
    unsigned long long off = old & ((1ull << kLiveShift) - 1ull);
    uint32_t li = (uint32_t)(old >> kLiveShift);
    for (int w = 0; w < 4; ++w) {
      const int id = wave_id(a, view, blockIdx.x, blockIdx.y, w);
      b.wave_off[id] = (unsigned long long)c << 56 | off;
      off += s_size[w];
      if (s_size[w]) b.live[(size_t)c * b.live_cap + li++] = id;
    }
  }
}

// Wave i of the waves with samples.
__device__ __forceinline__ int live_wave(const DiffBins &b, uint32_t i) {
  int c = 0;
  while (i >= b.live_pre[c + 1]) ++c;
  return b.live[(size_t)c * b.live_cap + (i - b.live_pre[c])];
}

// Resample pass of the adjoint: every sample the forward took (the first
// nsamples lattice points of the pixel's brick visits) is evaluated once
// more -- value tap and gradient, the seven trilinear taps -- and stored in
// its slot (sample k of N of the ray in tile place l: backward step N-1-k).  Nothing here
// depends on the compositing recurrence (a sample's index in the ray fixes
// it), so the G lanes of a group evaluate G consecutive lattice samples at
// once and a ray is done after its nsamples samples; the reverse walk then
// reads 16 bytes per sample instead of gathering its 20 footprint voxels.
// The brick walk is the forward's (closest reported hit from tmin, ties to
// the lower slot; jitter per visit), so the samples are the forward's bit
// for bit.  Each wave owns one 4x4 tile of an 8x8 block (few rays per wave:
// the rays with samples are a small part of the image, and more waves hide
// the gathers' latency better); a group takes the tile's next pixel with
// samples as soon as its ray is done.  Launched with 8 lanes per ray (8
// groups take the 16 pixels: 141 against 160 us with 16 groups of 4 on
// 256^3 x 8 views of 512^2).
template <int G>
__global__ __launch_bounds__(kBlock) void diff_resample_kernel(const DiffArgs a, const DiffBins b) {
  enum : int { kIdle = 0, kPick = 1, kMarch = 2, kOver = 3 };
  // block i: the 8x8 tile of live wave i
  int view, bx, by, tw;
  wave_of(a, live_wave(b, blockIdx.x), view, bx, by, tw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gl = lane % G, gbase = lane - gl;
  const int tx = bx * kTile + (tw & 1) * 8 + (wave & 1) * 4;
  const int ty = by * kTile + (tw >> 1) * 8 + (wave >> 1) * 4;
  const float *org = a.cam[view];
  const float dt = a.dt;
  int state = kIdle, next = 0;  // next: wave-uniform
  float d[3] = {1.f, 0.f, 0.f};
  uint32_t N = 0, done = 0, rng = 0;
  float4 *out_s = b.samples;  // the slot of the ray's last sample
  float tmin = 0.f, t1 = 0.f, j = 0.f, k0 = 0.f;
  int kk = 0;
  BrickView bv{nullptr, 0u, 0u, 0, 0, 0, nullptr, false};
  float lo0 = 0.f, lo1 = 0.f, lo2 = 0.f;
  while (__any(state != kOver)) {
    // a group without a ray takes the tile's next pixel
    const bool want = state == kIdle && gl == 0;
    const uint64_t m = __ballot(want);
    int k = next + __popcll(m & ((1ull << lane) - 1ull));
    next += __popcll(m);
    k = group_bcast_i<G, 0>(k, gbase);
    if (state == kIdle) {
      if (k >= 16) {
        state = kOver;
      } else {
        const int px = tx + (k & 3), py = ty + (k >> 2);
        if (px < a.W && py < a.H) {
          const size_t o = (size_t)view * a.W * a.H + px + (size_t)a.W * py;
          N = a.nsamples[o];
          if (N > 0) {
            ray_dir(a, view, px, py, d);
            rng = rng_seed((uint32_t)(px + a.W * py), 0u);
            out_s = b.samples + (ray_slot(a, b, view, px, py) + 64ull * (N - 1u));
            done = 0;
            tmin = 0.f;
            state = kPick;
          }
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
        state = kIdle;  // (only with sample counts that are not the forward's)
      } else {
        j = a.jitter_mode == ALOHA_JITTER_FIXED ? a.jitter_value : rng_next(rng);
        const float segs = floorf((bt0 - dt * j) / dt);
        const float tfirst = dt * (j + segs);
        k0 = tfirst < bt0 ? segs + 1.f : segs;
        t1 = bt1;
        kk = 0;
        const DevBrick &db = a.bricks[hit];
        bv = brick_view(db, nullptr);  // f32 bricks only (checked by the host)
        lo0 = db.vox_lo[0];
        lo1 = db.vox_lo[1];
        lo2 = db.vox_lo[2];
        tmin = bt1 + 1e-20f;
        state = kMarch;
      }
    }
    // kMarch: lattice samples kk .. kk + G - 1 of the segment; the lanes that
    // take one are a prefix of the group (t grows with the lane)
    const bool go = state == kMarch;
    const float t = dt * (j + (k0 + (float)(kk + gl)));
    const bool take = go && t < t1 && done + (uint32_t)gl < N;
    if (take) {
      const float x = (org[0] + t * d[0]) - lo0;
      const float y = (org[1] + t * d[1]) - lo1;
      const float z = (org[2] + t * d[2]) - lo2;
      Sample sm;
      if (!sample_fast<false>(bv, x, y, z, sm))
        sm = sample_exact<false, false>(bv, nullptr, a.bricks[0], x, y, z);
      out_s[-64 * (ptrdiff_t)(done + gl)] = make_float4(sm.val, sm.gx, sm.gy, sm.gz);
    }
    const uint64_t took = __ballot(take);
    if (go) {
      done += (uint32_t)__popcll((took >> gbase) & ((1ull << G) - 1ull));
      kk += G;
      if (done >= N) state = kIdle;
      else if (!(dt * (j + (k0 + (float)kk)) < t1)) state = kPick;
    }
  }
}

// Exclusive prefix of the bin counts (one workgroup): off[]; off[nbins] is
// the sample total.
__global__ __launch_bounds__(1024) void diff_scan_kernel(const DiffBins b) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x, n = b.nbins;
  const int per = (n + 1023) / 1024;
  const int i0 = min(n, t * per), i1 = min(n, i0 + per);
  uint32_t sum = 0;
  for (int i = i0; i < i1; ++i) sum += b.count[i];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (int i = i0; i < i1; ++i) {
    b.off[i] = run;
    run += b.count[i];
  }
  if (t == 1023) b.off[n] = part[1023];
}

// Runs into their bins: each head takes the next places of its bin for its
// samples and raises the bin's bounds.  Heads of one wave in the same bin
// (neighbouring rays entering it at the same backward step) are served by
// one atomic per bin: the leader adds their lengths, each head's rank is the
// leader's result plus the lengths of the heads before it.  The order of the
// runs inside a bin depends on the atomics; the bin's sums do not.
__global__ __launch_bounds__(kBlock) void diff_run_rank_kernel(const DiffBins b) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int bin = s < b.total ? b.run_bin[s] : -1;
  if (bin >= b.nbins) bin = -1;
  float4 info = make_float4(0.f, 0.f, 0.f, 0.f);
  if (bin >= 0) info = b.run_info[s];
  const uint32_t len = bin >= 0 ? __float_as_uint(info.x) : 0u;
  // bounds are >= 0 or NaN: as unsigned bits NaN is above every finite bound
  const uint32_t bd = __float_as_uint(info.y), bt = __float_as_uint(info.z);
  // Group the heads by bin (one trip per distinct bin, registers only),
  // then every group's first lane takes the group's places with one atomic
  // -- all of them in the same instruction, so a row waits for one atomic
  // round trip however many bins its heads fall into.
  bool pending = bin >= 0;
  int leader_of = 0;     // the first lane of this head's group
  uint32_t before = 0;   // the lengths of the group's heads before this one
  bool leads = false;    // this lane is its group's first
  uint32_t gsum = 0, gmd = 0, gmt = 0;  // (leaders) the group's length and bounds
  while (true) {  // wave-uniform
    const uint64_t pm = __ballot(pending);
    if (!pm) break;
    const int leader = __ffsll((unsigned long long)pm) - 1;
    const int lb = __shfl(bin, leader);
    const bool mine = pending && bin == lb;
    // inclusive scan of the lengths of this bin's heads, and their bounds' maxima
    uint32_t v = mine ? len : 0u, md = mine ? bd : 0u, mt = mine ? bt : 0u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)v, o);
      if (lane >= o) v += u;
      md = max(md, (uint32_t)__shfl_xor((int)md, o));
      mt = max(mt, (uint32_t)__shfl_xor((int)mt, o));
    }
    const uint32_t sum = (uint32_t)__shfl((int)v, 63);
    if (lane == leader) {
      leads = true;
      gsum = sum;
      gmd = md;
      gmt = mt;
    }
    if (mine) {
      leader_of = leader;
      before = v - len;
      pending = false;
    }
  }
  uint32_t base = 0;
  if (leads) {
    base = atomicAdd(&b.count[bin], gsum);
    if (gmd) atomicMax(&b.maxd[bin], gmd);
    if (gmt) atomicMax(&b.maxt[bin], gmt);
  }
  base = (uint32_t)__shfl((int)base, leader_of);
  if (bin >= 0) {
    info.w = __uint_as_float(base + before);
    b.run_info[s] = info;
  }
}

// The bins' slot lists: run (head s, length L, rank r) of bin c fills
// rec_list[off[c] + r ..] with s, s - 64, .., s - 64 (L - 1).
__global__ __launch_bounds__(kBlock) void diff_run_list_kernel(const DiffBins b) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= b.total) return;
  const int bin = b.run_bin[s];
  if (bin < 0 || bin >= b.nbins) return;
  const float4 info = b.run_info[s];
  const uint32_t len = __float_as_uint(info.x), base = b.off[bin] + __float_as_uint(info.w);
  const uint32_t end = min(base + len, b.off[bin + 1]);
  for (uint32_t i = base; i < end; ++i) b.rec_list[i] = s - 64u * (i - base);
}

// The adjoint march: per pixel, the forward's samples in reverse (T_k =
// T_{k+1}/(1-a_k)), back-propagated through compositing, shading and the TF
// lookup.  Each sample becomes a record in its slot: its position, the
// derivative of the loss with respect to its value tap and to its gradient
// (through N = g/|g+1e-4|), its value (which fixes the two TF entries it
// read) and the derivative with respect to the TF colour and opacity it got;
// diff_bin_kernel scatters them.  A sample's value tap and gradient come
// from the resample pass (16 bytes, read one sample ahead), so the walk
// issues no voxel gathers; its bin runs are marked at their head slots (the
// bin, the length and the largest record bounds), so it needs no atomics
// either -- diff_run_rank_kernel and diff_run_list_kernel place the runs.
//
// One lane per ray, 8x8 pixels per wave.  Measured on 256^3 x 8 views of
// 512^2: 1.36 ms while the walk evaluated the taps itself and reserved
// records in bin order (an atomic with return at each bin run, whose
// length it counted by stepping through the run first): each loop trip
// waited for some lane's atomic and run count; without the run count 0.67
// ms, without the atomics as well 0.20 ms.  A lane group per ray (repeating
// the set-up on every lane of the group) 1.53 ms (4 lanes); persistent lanes
// taking pixels of a 16x16 region per wave as their rays end, 2.5 ms.
__global__ __launch_bounds__(kBlock) void diff_backward_kernel(const DiffArgs a, const DiffBins b) {
  extern __shared__ float4 s_tf[];
  load_tf(a, s_tf);
  __syncthreads();

  // wave i of the launch walks live wave i
  const int lane = threadIdx.x & 63;
  const uint32_t li = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (li >= b.live_pre[kSlotCounters]) return;
  int view, bx, by, wave;
  wave_of(a, live_wave(b, li), view, bx, by, wave);
  const int px = bx * kTile + (wave & 1) * 8 + (lane & 7);
  const int py = by * kTile + (wave >> 1) * 8 + (lane >> 3);
  if (px >= a.W || py >= a.H) return;
  const size_t out = (size_t)view * a.W * a.H + (px + a.W * py);
  const uint32_t N = a.nsamples[out];
  if (N == 0) return;
  const int pixel = px + a.W * py;
  const float *org = a.cam[view];
  float d[3];
  ray_dir(a, view, px, py, d);
  const float4 C = a.rgba[out];
  const float4 G = a.dldc[out];
  // the visit holding the last sample, and the samples before it
  int last = -1;
  uint32_t before_last = 0, acc = 0;
  for (int vi = 0; vi < kMaxVisits && last < 0; ++vi) {
    Visit v;
    if (visits(a, org, d, (uint32_t)pixel, vi + 1, &v, vi) <= vi) break;
    const uint32_t cnt = visit_count(a, v, N - acc);
    if (acc + cnt >= N) {
      last = vi;
      before_last = acc;
    }
    acc += cnt;
  }
  // the ray's slots: backward step i at slot0 + 64 i
  const uint32_t slot0 = (uint32_t)ray_slot(a, b, view, px, py);
  const float4 *smp = b.samples + slot0;
  uint32_t idx = N;  // samples not yet walked back
  float4 nxt = smp[0];
  float T = 1.f - C.w;  // transmittance after the last sample
  const float TN = T;
  float Sr = 0.f, Sg = 0.f, Sb = 0.f;  // suffix sums of later contributions
  // the current bin run (its lowest slot so far is its head) and the
  // largest bounds of its records
  int run_bin = -1;
  uint32_t run_len = 0, head = 0;
  float rmd = 0.f, rmt = 0.f;
  auto close_run = [&]() {
    if (run_bin >= 0) {
      b.run_bin[head] = run_bin;
      b.run_info[head] = make_float4(__uint_as_float(run_len), rmd, rmt, 0.f);
    }
  };
  const int n_tf = a.tf_n;
  for (int vi = last; vi >= 0; --vi) {
    Visit v;
    visits(a, org, d, (uint32_t)pixel, vi + 1, &v, vi);
    const uint32_t cnt = vi == last ? N - before_last : visit_count(a, v, 0xFFFFFFFFu);
    // The visit's brick and bin table, read here and waited for once: the
    // vector memory counter also counts stores, so a wait for these inside
    // the loop below (where the compiler cannot tell that they arrived in an
    // earlier step) would make every step wait for the previous step's
    // record stores and for the sample it has just requested.  The empty asm
    // takes them as its operands, so the loop uses its results instead of
    // the loads.
    const DevBrick &db = a.bricks[v.bid];
    float lo0 = db.vox_lo[0], lo1 = db.vox_lo[1], lo2 = db.vox_lo[2];
    int n0 = db.n[0], n1 = db.n[1], n2 = db.n[2];
    int bbase = b.base[v.slot], bnx = b.nbx[v.slot], bny = b.nby[v.slot];
    asm volatile("" : "+v"(lo0), "+v"(lo1), "+v"(lo2), "+v"(n0), "+v"(n1), "+v"(n2), "+v"(bbase),
                 "+v"(bnx), "+v"(bny));
    for (int k = (int)cnt - 1; k >= 0 && idx > 0; --k) {
      const float t = lattice_t(a, v, k);
      Sample sv;
      sv.val = nxt.x;
      sv.gx = nxt.y;
      sv.gy = nxt.z;
      sv.gz = nxt.w;
      --idx;
      if (idx > 0) nxt = smp[64 * (N - idx)];
      const Shaded r = shade_of(a, s_tf, sv, d);
      const float am = r.tf.w;
      const float one_m = 1.f - am;
      const float inv_om = 1.f / one_m;
      const float Tk = T * inv_om;
      const float cs_r = r.tf.x * r.shade, cs_g = r.tf.y * r.shade, cs_b = r.tf.z * r.shade;
      const float Gc = G.x * r.tf.x + G.y * r.tf.y + G.z * r.tf.z;
      // d/d(colour), d/d(shade), d/d(opacity)
      const float Ta = Tk * am;
      const float dcr = G.x * Ta * r.shade, dcg = G.y * Ta * r.shade, dcb = G.z * Ta * r.shade;
      const float dshade = Ta * Gc;
      const float dalpha = Tk * (G.x * cs_r + G.y * cs_g + G.z * cs_b) -
                           ((G.x * Sr + G.y * Sg + G.z * Sb) - G.w * TN) * inv_om;
      Sr += Ta * cs_r;
      Sg += Ta * cs_g;
      Sb += Ta * cs_b;
      T = Tk;
      const float da = r.raw_a < a.alpha_max ? dalpha : 0.f;
      // value tap through the TF slope (the entries tf_lookup read)
      const float xt = r.s.val * (float)n_tf - 0.5f;
      const int i0 = (int)fminf(fmaxf(floorf(xt), -2.f), (float)n_tf);
      const float4 e0 = s_tf[i0 + 2], e1 = s_tf[i0 + 3];
      const float dv = (dcr * (e1.x - e0.x) + dcg * (e1.y - e0.y) + dcb * (e1.z - e0.z) +
                        da * (e1.w - e0.w)) * (float)n_tf;
      // shading: s = .1 + .9|N.d|, N = g/L, L = |g + 1e-4|; the side of the
      // |N.d| kink as the forward took it
      const float nd = r.nd;
      const float sg = nd > 0.f ? .9f : (nd < 0.f ? -.9f : 0.f);
      const float dN[3] = {dshade * sg * d[0], dshade * sg * d[1], dshade * sg * d[2]};
      const float gv[3] = {r.s.gx, r.s.gy, r.s.gz};
      const float dNg = dN[0] * gv[0] + dN[1] * gv[1] + dN[2] * gv[2];
      const float il = 1.f / r.len, il3 = il * il * il;
      float dg[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) dg[c] = dN[c] * il - dNg * (gv[c] + 1e-4f) * il3;
      const float x = (org[0] + t * d[0]) - lo0;  // the sample's brick-texel position
      const float y = (org[1] + t * d[1]) - lo1;
      const float z = (org[2] + t * d[2]) - lo2;
      const int bin = bin_of(bbase, bnx, bny, n0, n1, n2, x, y, z);
      if (bin != run_bin) {  // the run walked so far ended at its head
        close_run();
        run_bin = bin;
        run_len = 0;
        rmd = rmt = 0.f;
      }
      const uint32_t sl = slot0 + 64u * (N - 1u - idx);
      head = sl;
      ++run_len;
      {
        float4 *o = b.records + (size_t)kRecWords * sl;
        o[0] = make_float4(x, y, z, dv);
        o[1] = make_float4(dg[0], dg[1], dg[2], r.s.val);
        o[2] = make_float4(dcr, dcg, dcb, da);
        // bounds of the record's contributions: density |dv| + 2 (|dgx| +
        // |dgy| + |dgz|) (the weights of a tap sum to 1), TF the largest of
        // |dc|, |da|; NaN wins (as unsigned bits it is above every finite
        // bound and infinity)
        const float bd = fabsf(dv) + 2.f * (fabsf(dg[0]) + fabsf(dg[1]) + fabsf(dg[2]));
        // (a NaN must win here as well: fmaxf would drop it, and the bin's
        // fixed-point scale would then turn the NaN into an integer)
        const float a0 = fabsf(dcr), a1 = fabsf(dcg), a2 = fabsf(dcb), a3 = fabsf(da);
        const float m01 = (a0 > a1 || a0 != a0) ? a0 : a1, m23 = (a2 > a3 || a2 != a2) ? a2 : a3;
        const float bt = (m01 > m23 || m01 != m01) ? m01 : m23;
        rmd = (bd > rmd || bd != bd) ? bd : rmd;
        rmt = (bt > rmt || bt != bt) ? bt : rmt;
      }
    }
  }
  close_run();
}

// Scale 2^S of a bin's fixed-point sums: at most `count` contributions of at
// most `bound` each stay below 2^61 in magnitude.  0 when there is nothing
// to add, -1 when the bound is not finite (then the sums are NaN).
__device__ __forceinline__ double fixed_scale(float bound, uint32_t count) {
  if (!(bound <= 3.0e38f)) return -1.0;
  if (bound == 0.f) return 0.0;
  int e;
  frexp((double)bound * (double)count, &e);  // bound * count < 2^e
  return ldexp(1.0, 61 - e);
}

// Adds a run's fixed-point TF sums (entries i and i + 1 of the chunk
// [e0, e1); i == INT_MIN: no run) to the LDS sums.
__device__ __forceinline__ void tf_run_flush(unsigned long long *tfs, int i, int e0, int e1,
                                             const long long q0[4], const long long q1[4]) {
  if (i == INT_MIN) return;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (i >= e0 && i < e1 && q0[c] != 0)
      __hip_atomic_fetch_add((lds_u64 *)tfs + 4 * (i - e0) + c, (unsigned long long)q0[c],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (i + 1 >= e0 && i + 1 < e1 && q1[c] != 0)
      __hip_atomic_fetch_add((lds_u64 *)tfs + 4 * (i + 1 - e0) + c, (unsigned long long)q1[c],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// Threads of a bin's workgroup: its LDS (the 18^3 tile and the TF sums,
// ~54 KB) allows two workgroups per CU, so a 512-thread workgroup keeps 4
// waves on each SIMD (256 threads: 2).
constexpr int kBinBlock = 512;

// One bin: its records' density gradient scattered onto an 18^3 tile (each
// sample's 20 footprint voxels, or the 56 corners of its seven taps near a
// brick face) and their TF gradient onto the TF entries, both as 64-bit
// fixed-point LDS sums; then the tile goes out as floats to b.tiles and the
// TF sums as doubles to b.tf_part.  Order-independent sums: bit-reproducible.
__global__ __launch_bounds__(kBinBlock) void diff_bin_kernel(const DiffArgs a, const DiffBins b) {
  extern __shared__ unsigned long long s_fix[];
  unsigned long long *tile = s_fix;               // kTileVox
  unsigned long long *tfs = s_fix + kTileVox;     // 4 x min(tf_n, kLdsTfChunk)
  const int bin = blockIdx.x;
  const uint32_t r0 = b.off[bin], r1 = b.off[bin + 1];
  if (r0 >= r1) return;  // block-uniform
  int slot = 0;  // the brick slot holding this bin
  while (b.base[slot + 1] <= bin) ++slot;
  const DevBrick db = a.bricks[a.brick_ids[slot]];
  const bool dens = a.dvox[slot] != nullptr;
  const bool want_tf = a.dtf != nullptr;
  const int n = a.tf_n;
  // a bin's slots come in runs of one ray's consecutive samples (half a
  // voxel apart, footprints overlapping): each thread takes a contiguous
  // chunk of the list, so the lanes of one instruction rarely add to the
  // same word
  const uint32_t per = (r1 - r0 + kBinBlock - 1) / kBinBlock;
  const uint32_t c0 = r0 + per * threadIdx.x, c1 = min(r1, c0 + per);
  // the largest bounds of the bin's records (raised by the adjoint march)
  const float md = __uint_as_float(b.maxd[bin]), mt = __uint_as_float(b.maxt[bin]);
  const uint32_t cnt = r1 - r0;
  const double sd = dens ? fixed_scale(md, 8u * cnt) : 0.0;  // 8 corners of a tap share a voxel
  const double st = want_tf ? fixed_scale(mt, cnt) : 0.0;
  const int l = bin - b.base[slot], nbx = b.nbx[slot], nby = b.nby[slot];
  const int bx = l % nbx, by = (l / nbx) % nby, bz = l / (nbx * nby);
  const FixedTile sink{tile, bx * kBinEdge - 1, by * kBinEdge - 1, bz * kBinEdge - 1, sd};
  const int chunks = want_tf ? (n + kLdsTfChunk - 1) / kLdsTfChunk : 1;
  for (int ch = 0; ch < chunks; ++ch) {
    const int e0 = ch * kLdsTfChunk, e1 = min(n, e0 + kLdsTfChunk);
    const bool do_dens = ch == 0 && sd > 0.0;
    if (do_dens)
      for (int i = threadIdx.x; i < kTileVox; i += kBinBlock) tile[i] = 0ull;
    for (int i = threadIdx.x; i < 4 * (e1 - e0); i += kBinBlock) tfs[i] = 0ull;
    __syncthreads();
    int ri = INT_MIN;  // the TF entry pair of the current run, its sums
    long long q0[4] = {0, 0, 0, 0}, q1[4] = {0, 0, 0, 0};
    // The thread's records are loaded one ahead of their scatter and its
    // list entries two ahead, so the loads of the next record are in flight
    // while this one's LDS sums run (each record is its own cache lines:
    // consecutive records of a run lie 64 slots apart).  A slot at or past
    // b.total (only with sample counts that are not the forward's) is
    // skipped.
    const bool want_t = st > 0.0;
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
    uint32_t nsl = c0 < c1 ? b.rec_list[c0] : b.total;
    uint32_t nsl2 = c0 + 1 < c1 ? b.rec_list[c0 + 1] : b.total;
    auto fetch = [&](uint32_t sl) {
      if (sl < b.total) {
        const float4 *p = b.records + (size_t)kRecWords * sl;
        n0 = p[0];
        n1 = p[1];
        if (want_t) n2 = p[2];
      }
    };
    fetch(nsl);
    for (uint32_t r = c0; r < c1; ++r) {
      const uint32_t sl = nsl;
      const float4 p0 = n0, p1 = n1, p2 = n2;
      nsl = nsl2;
      if (r + 2 < c1) nsl2 = b.rec_list[r + 2];
      if (r + 1 < c1) fetch(nsl);
      if (sl >= b.total) continue;
      if (do_dens) {
        const float gd[3] = {p1.x, p1.y, p1.z};
        if (p0.w != 0.f || gd[0] != 0.f || gd[1] != 0.f || gd[2] != 0.f) {
          if (!scatter_fast(sink, db, p0.x, p0.y, p0.z, p0.w, gd)) {
            scatter_tap(sink, db, p0.x, p0.y, p0.z, p0.w);
            scatter_tap(sink, db, p0.x + .5f, p0.y, p0.z, gd[0]);
            scatter_tap(sink, db, p0.x - .5f, p0.y, p0.z, -gd[0]);
            scatter_tap(sink, db, p0.x, p0.y + .5f, p0.z, gd[1]);
            scatter_tap(sink, db, p0.x, p0.y - .5f, p0.z, -gd[1]);
            scatter_tap(sink, db, p0.x, p0.y, p0.z + .5f, gd[2]);
            scatter_tap(sink, db, p0.x, p0.y, p0.z - .5f, -gd[2]);
          }
        }
      }
      if (st > 0.0) {
        // the two TF entries of the lookup (tf_lookup's texel and weight);
        // consecutive records of a thread are neighbouring samples of one
        // ray and mostly read the same entry pair: their fixed-point values
        // are summed in registers and added to LDS once per run
        const float4 g4 = p2;
        const float xt = p1.w * (float)n - 0.5f;
        const float fi = floorf(xt);
        const float w = xt - fi;
        const int i0 = (int)fminf(fmaxf(fi, -2.f), (float)n);
        if (i0 != ri) {
          tf_run_flush(tfs, ri, e0, e1, q0, q1);
          ri = i0;
#pragma unroll
          for (int c = 0; c < 4; ++c) q0[c] = q1[c] = 0;
        }
        const float gs[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          q0[c] += to_fixed((1.f - w) * gs[c], st);
          q1[c] += to_fixed(w * gs[c], st);
        }
      }
    }
    tf_run_flush(tfs, ri, e0, e1, q0, q1);
    __syncthreads();
    if (ch == 0 && dens) {
      float *dst = b.tiles + (size_t)bin * kTileVox;
      const double inv = sd > 0.0 ? 1.0 / sd : 0.0;
      for (int i = threadIdx.x; i < kTileVox; i += kBinBlock)
        dst[i] = sd > 0.0 ? (float)((double)(long long)tile[i] * inv)
                          : (sd < 0.0 ? __int_as_float(0x7fc00000) : 0.f);
    }
    if (want_tf) {
      double *dst = b.tf_part + (size_t)bin * 4 * n;
      const double inv = st > 0.0 ? 1.0 / st : 0.0;
      for (int i = 4 * e0 + threadIdx.x; i < 4 * e1; i += kBinBlock)
        dst[i] = st > 0.0 ? (double)(long long)tfs[i - 4 * e0] * inv
                          : (st < 0.0 ? __longlong_as_double(0x7ff8000000000000ll) : 0.0);
    }
    __syncthreads();
  }
}

// The TF gradient in two fixed-order levels: group g of kTfGroupBins
// consecutive bins sums its bins' partials in bin order (doubles), then each
// entry sums the groups in order and adds the result into dtf.
__global__ __launch_bounds__(kBlock) void diff_tf_group_kernel(const DiffArgs a, const DiffBins b,
                                                               double *groups) {
  const int i = blockIdx.x * kBlock + threadIdx.x, g = blockIdx.y;
  const int m = 4 * a.tf_n;
  if (i >= m) return;
  const int b0 = g * kTfGroupBins, b1 = min(b.nbins, b0 + kTfGroupBins);
  double s = 0.0;
#pragma unroll 8
  for (int bin = b0; bin < b1; ++bin)  // (loads issued ahead, added in bin order)
    if (b.off[bin] != b.off[bin + 1]) s += b.tf_part[(size_t)bin * m + i];
  groups[(size_t)g * m + i] = s;
}

__global__ __launch_bounds__(kBlock) void diff_tf_reduce_kernel(const DiffArgs a, const DiffBins b,
                                                                const double *groups, int ngroups) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int m = 4 * a.tf_n;
  if (i >= m) return;
  double s = 0.0;
#pragma unroll 8
  for (int g = 0; g < ngroups; ++g) s += groups[(size_t)g * m + i];  // (in group order)
  a.dtf[i] += (float)s;
}

// Tiles into the gradient grid, one workgroup per bin: voxel v of an axis
// lies in the tile of block (v + 1) / 16 (at l = (v + 1) % 16) and, when
// l < 2, also in the previous block's (at l + 16): up to 8 tiles, added in a
// fixed order (per axis the previous block's first).  The workgroup of
// bin b folds the 16^3 voxels with (v + 1) / 16 == b per axis; the 8 tiles
// that can reach them are known per workgroup, and a tile of a bin without
// samples was not written and is skipped.  Each thread folds one z column
// of 16 voxels: the tiles' and the grid's x rows are read contiguously.
__global__ __launch_bounds__(kBlock) void diff_fold_kernel(const DiffArgs a, const DiffBins b) {
  static_assert(kBlock == kBinEdge * kBinEdge, "a thread per (x, y) of the bin");
  const int bin = blockIdx.x;
  int slot = 0;  // the brick slot holding this bin (block-uniform)
  while (b.base[slot + 1] <= bin) ++slot;
  float *grad = a.dvox[slot];
  if (!grad) return;
  const int nbx = b.nbx[slot], nby = b.nby[slot];
  const int l = bin - b.base[slot];
  const int bx = l % nbx, by = (l / nbx) % nby, bz = l / (nbx * nby);
  // has[dz][dy][dx]: the tile of block (bz - dz, by - dy, bx - dx) exists and was written
  bool has[2][2][2], any = false;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const bool in = bz >= dz && by >= dy && bx >= dx;
        has[dz][dy][dx] = in && b.count[bin - (dz * nby + dy) * nbx - dx] != 0;
        any |= has[dz][dy][dx];
      }
  if (!any) return;
  const DevBrick &db = a.bricks[a.brick_ids[slot]];
  const int nx = db.n[0], ny = db.n[1], nz = db.n[2];
  const int lx = threadIdx.x % kBinEdge, ly = threadIdx.x / kBinEdge;
  const int x = bx * kBinEdge - 1 + lx, y = by * kBinEdge - 1 + ly;
  if (x < 0 || x >= nx || y < 0 || y >= ny) return;
  const int dxs = lx < 2 ? 1 : 0, dys = ly < 2 ? 1 : 0;  // the previous block's tile reaches x / y
  for (int lz = 0; lz < kBinEdge; ++lz) {
    const int z = bz * kBinEdge - 1 + lz;
    if (z < 0 || z >= nz) continue;
    float sum = 0.f;
    for (int dz = lz < 2 ? 1 : 0; dz >= 0; --dz)
      for (int dy = dys; dy >= 0; --dy)
        for (int dx = dxs; dx >= 0; --dx) {
          if (!has[dz][dy][dx]) continue;
          const int tb = bin - (dz * nby + dy) * nbx - dx;
          sum += b.tiles[(size_t)tb * kTileVox +
                         ((lz + kBinEdge * dz) * kTileEdge + (ly + kBinEdge * dy)) * kTileEdge +
                         (lx + kBinEdge * dx)];
        }
    if (sum != 0.f) {
      float *gp = grad + ((size_t)z * ny + y) * nx + x;
      *gp = *gp + sum;
    }
  }
}


}  // namespace

hipError_t launch_diff(const DiffArgs &a, bool backward, hipStream_t s) {
  if (backward) return hipErrorInvalidValue;  // the backward runs in stages (launch_diff_*)
  const int fill_blocks = std::min(a.V * a.H, 1024);
  diff_forward_group_kernel<4><<<a.block_off[a.V] + fill_blocks, kBlock,
                                 sizeof(float4) * (size_t)(a.tf_n + 4), s>>>(a);
  return hipGetLastError();
}

hipError_t launch_diff_rays(const DiffArgs &a, const DiffBins &b, hipStream_t s) {
  const dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, a.V);
  diff_wave_slots_kernel<<<grid, kBlock, 0, s>>>(a, b);
  return hipGetLastError();
}

hipError_t launch_diff_backward(const DiffArgs &a, const DiffBins &b, hipStream_t s) {
  const uint32_t nlive = b.live_pre[kSlotCounters];
  if (nlive > 0) {  // (no live wave: no slots either)
    diff_resample_kernel<8><<<nlive, kBlock, 0, s>>>(a, b);
    diff_backward_kernel<<<(nlive + kBlock / 64 - 1) / (kBlock / 64), kBlock,
                           sizeof(float4) * (size_t)(a.tf_n + 4), s>>>(a, b);
  }
  if (b.total > 0) {
    const unsigned blocks = (unsigned)((b.total + kBlock - 1) / kBlock);
    diff_run_rank_kernel<<<blocks, kBlock, 0, s>>>(b);
    diff_scan_kernel<<<1, 1024, 0, s>>>(b);
    diff_run_list_kernel<<<blocks, kBlock, 0, s>>>(b);
  } else {
    diff_scan_kernel<<<1, 1024, 0, s>>>(b);
  }
  return hipGetLastError();
}

hipError_t launch_diff_bins(const DiffArgs &a, const DiffBins &b, const int max_dims[3],
                            hipStream_t s) {
  if (b.nbins == 0) return hipSuccess;
  const size_t lds =
      sizeof(unsigned long long) * ((size_t)kTileVox + 4 * (size_t)std::min(a.tf_n, kLdsTfChunk));
  diff_bin_kernel<<<b.nbins, kBinBlock, lds, s>>>(a, b);
  if (max_dims[0] > 0) diff_fold_kernel<<<b.nbins, kBlock, 0, s>>>(a, b);
  if (a.dtf) {
    // the group sums go behind the bins' partials in b.tf_part
    const int m = 4 * a.tf_n, ngroups = (b.nbins + kTfGroupBins - 1) / kTfGroupBins;
    double *groups = b.tf_part + (size_t)b.nbins * m;
    const unsigned bx = (unsigned)((m + kBlock - 1) / kBlock);
    diff_tf_group_kernel<<<dim3(bx, (unsigned)ngroups), kBlock, 0, s>>>(a, b, groups);
    diff_tf_reduce_kernel<<<bx, kBlock, 0, s>>>(a, b, groups, ngroups);
  }
  return hipGetLastError();
}

}  // namespace aloha
