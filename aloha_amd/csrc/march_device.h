// march_device.h -- device building blocks of the ray march shared by the
// plain raygen kernel (march.hip) and the differentiable forward/adjoint
// kernels (adjoint.hip): RNG, brick addressing, the exact trilinear tap, the
// shared 20-voxel gradient footprint, the LDS transfer-function lookup and
// the slab test.  All arithmetic is written in the CPU restatement's order
// and must be compiled without contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "device_types.h"

namespace aloha {
namespace dev {

__device__ __forceinline__ float lerp_(float a, float b, float w) {
  return __builtin_fmaf(w, b - a, a);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// --- RNG: owl::common::LCG<4> (deviceCode.cpp:12, :354, :386) --------------
// The published construction (Zafar, Olano & Curtis, HPG 2010): 4 rounds of
// TEA over (pixelID, accumID) with the key {a341316c, c8013ea4, ad90777d,
// 7e95761e} give the state; each draw steps the Numerical-Recipes LCG and
// returns its low 24 bits / 2^24.
__device__ __forceinline__ uint32_t rng_seed(uint32_t pixel, uint32_t frame) {
  uint32_t v0 = pixel, v1 = frame, sum = 0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
__device__ __forceinline__ float rng_next(uint32_t &s) {
  s = s * 1664525u + 1013904223u;
  return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// Voxel reads go through global-address-space pointers, so the compiler
// emits global loads (vector-memory counter only) rather than flat loads,
// which also count against the LDS counter and so would serialize with the
// TF lookups and lane shuffles.
typedef const float __attribute__((address_space(1))) gfloat;
typedef const uint8_t __attribute__((address_space(1))) gbyte;

// o points at padded voxel (-1,-1,-1), so every index the taps form is
// non-negative after +1 and all offsets are 24-bit products.  u8: the brick
// stores u8 voxels, decoded through the LDS table lut[u] = u/255.f (the
// float Volume::get gives, volume.h:78, bit for bit); lut is the launch's
// table (null in kernels compiled without the u8 decode).
struct BrickView {
  const gfloat *o;  // f32 voxels, or the u8 voxels' bytes
  uint32_t px, pxy;
  int n0, n1, n2;
  const float *lut;
  bool u8;
};

__device__ __forceinline__ BrickView brick_view(const DevBrick &db, const float *lut) {
  return BrickView{(const gfloat *)db.origin, (uint32_t)db.px, (uint32_t)db.pxy, db.n[0], db.n[1],
                   db.n[2], lut, lut != nullptr && db.dtype == ALOHA_U8};
}

// N consecutive voxels of one x-row from element e, in kernels that handle
// u8 bricks too: the same instructions for both storages (no branch, so no
// second gather holds registers): N dwords from byte e of a u8 brick (an unaligned load; the allocation's slack covers the bytes past
// the grid) or from element e of an f32 brick, then per voxel the decoded
// byte or the float, by the brick's storage.
template <int N, bool kU8>
__device__ __forceinline__ void read_row(const BrickView &b, uint32_t e, float *v) {
  static_assert(kU8, "f32-only kernels read their rows directly");
  const gbyte *p = (const gbyte *)b.o + (b.u8 ? e : 4u * e);
  uint32_t w[N];
  __builtin_memcpy(w, (const void *)p, 4 * N);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float dec = b.lut[(w[0] >> (8 * i)) & 0xFFu];
    v[i] = b.u8 ? dec : __uint_as_float(w[i]);
  }
}


// The decode table: lut[u] = u/255.f (IEEE division), filled by the block.
__device__ __forceinline__ void load_u8_lut(float *lut) {
  for (int i = threadIdx.x; i < kU8Lut; i += blockDim.x) lut[i] = (float)i / 255.f;
}

__device__ __forceinline__ uint32_t voff(const BrickView &b, int i, int j, int k) {
  return (uint32_t)(i + 1) + __umul24((uint32_t)(j + 1), b.px) + __umul24((uint32_t)(k + 1), b.pxy);
}

// Footprint pass: mark the clamp-to-edge voxel that a corner read hits.
__device__ __forceinline__ void mark(uint32_t *touched, const DevBrick &db, int x, int y,
                                     int z) {
  x = clampi(x, 0, db.n[0] - 1);
  y = clampi(y, 0, db.n[1] - 1);
  z = clampi(z, 0, db.n[2] - 1);
  const uint64_t idx = ((uint64_t)z * db.n[1] + y) * db.n[0] + x;
  atomicOr(&touched[db.touched_word + (idx >> 5)], 1u << (idx & 31));
}

// One trilinear fetch, exactly as the restated texture unit does it.  kU8:
// the decode of u8 voxels is compiled in (sample_exact's and sample_fast's
// kU8 likewise); without it the brick must hold f32 voxels.
template <bool kFoot, bool kU8>
__device__ __forceinline__ float tap_exact(const BrickView &b, uint32_t *touched,
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
        for (int i = 0; i < 2; ++i) mark(touched, db, ix + i, iy + j, iz + k);
  }
  // 32-bit element indices per corner row: no 64-bit address offsets stay
  // live across a caller's loop for this (rarely taken) path
  const uint32_t i00 = voff(b, ix, iy, iz), i10 = i00 + b.px, i01 = i00 + b.pxy,
                 i11 = i10 + b.pxy;
  float v[8];
  if constexpr (kU8) {
    read_row<2, kU8>(b, i00, v);
    read_row<2, kU8>(b, i10, v + 2);
    read_row<2, kU8>(b, i01, v + 4);
    read_row<2, kU8>(b, i11, v + 6);
  } else {
    v[0] = b.o[i00], v[1] = b.o[i00 + 1u], v[2] = b.o[i10], v[3] = b.o[i10 + 1u];
    v[4] = b.o[i01], v[5] = b.o[i01 + 1u], v[6] = b.o[i11], v[7] = b.o[i11 + 1u];
  }
  const float c00 = lerp_(v[0], v[1], wx);
  const float c10 = lerp_(v[2], v[3], wx);
  const float c01 = lerp_(v[4], v[5], wx);
  const float c11 = lerp_(v[6], v[7], wx);
  const float c0 = lerp_(c00, c10, wy);
  const float c1 = lerp_(c01, c11, wy);
  return lerp_(c0, c1, wz);
}

struct Sample {
  float val, gx, gy, gz;
};

template <bool kFoot, bool kU8>
__device__ __forceinline__ Sample sample_exact(const BrickView &b, uint32_t *touched,
                                               const DevBrick &db, float x, float y, float z) {
  // one tap at a time (scheduling barriers between them): this path is
  // rare, and keeping its 56 loads in flight at once would raise the
  // register count of every kernel that inlines it
  Sample s;
  const float xp = tap_exact<kFoot, kU8>(b, touched, db, x + .5f, y, z);
  __builtin_amdgcn_sched_barrier(0);
  const float xm = tap_exact<kFoot, kU8>(b, touched, db, x - .5f, y, z);
  __builtin_amdgcn_sched_barrier(0);
  const float yp = tap_exact<kFoot, kU8>(b, touched, db, x, y + .5f, z);
  __builtin_amdgcn_sched_barrier(0);
  const float ym = tap_exact<kFoot, kU8>(b, touched, db, x, y - .5f, z);
  __builtin_amdgcn_sched_barrier(0);
  const float zp = tap_exact<kFoot, kU8>(b, touched, db, x, y, z + .5f);
  __builtin_amdgcn_sched_barrier(0);
  const float zm = tap_exact<kFoot, kU8>(b, touched, db, x, y, z - .5f);
  __builtin_amdgcn_sched_barrier(0);
  s.val = tap_exact<kFoot, kU8>(b, touched, db, x, y, z);
  s.gx = xp - xm;
  s.gy = yp - ym;
  s.gz = zp - zm;
  return s;
}

// Shared 20-voxel footprint.  Returns false (and computes nothing) when the
// lane's taps do not fit the window; the caller then uses sample_exact.
template <bool kU8>
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

  // R[jj][kk][i] = V(ax+i, cy+jj, cz+kk): four x-rows of 3 voxels; YN / ZN:
  // the y row outside the value tap's pair, and the z row likewise
  const int yn = scy ? ay : ay + 2;
  const int zn = scz ? az : az + 2;
  float R[2][2][3], YN[2][2], ZN[2][2];
  if (kU8 && b.u8) {
    // u8 voxels: each row's bytes in one unaligned load (4 bytes for the
    // 3-voxel rows, 2 for the pairs; the allocation has slack past its
    // last row), decoded through the table
    const gbyte *o8 = (const gbyte *)b.o;
    const float *lut = b.lut;
    auto ld4 = [&](uint32_t e) {
      uint32_t w;
      __builtin_memcpy(&w, (const void *)(o8 + e), 4);
      return w;
    };
    auto ld2 = [&](uint32_t e) {
      uint16_t w;
      __builtin_memcpy(&w, (const void *)(o8 + e), 2);
      return (uint32_t)w;
    };
    const uint32_t e00 = voff(b, ax, cy, cz);
    const uint32_t wr[2][2] = {{ld4(e00), ld4(e00 + pxy)}, {ld4(e00 + px), ld4(e00 + px + pxy)}};
    const uint32_t ey = voff(b, cx, yn, cz), ez = voff(b, cx, cy, zn);
    const uint32_t wy0 = ld2(ey), wy1 = ld2(ey + pxy), wz0 = ld2(ez), wz1 = ld2(ez + px);
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int i = 0; i < 3; ++i) R[jj][kk][i] = lut[(wr[jj][kk] >> (8 * i)) & 0xFFu];
    YN[0][0] = lut[wy0 & 0xFFu];
    YN[0][1] = lut[wy0 >> 8];
    YN[1][0] = lut[wy1 & 0xFFu];
    YN[1][1] = lut[wy1 >> 8];
    ZN[0][0] = lut[wz0 & 0xFFu];
    ZN[0][1] = lut[wz0 >> 8];
    ZN[1][0] = lut[wz1 & 0xFFu];
    ZN[1][1] = lut[wz1 >> 8];
  } else {  // f32 voxels: one row pointer per row, the x offsets immediates
    const gfloat *r00 = b.o + voff(b, ax, cy, cz);
    const gfloat *r01 = r00 + pxy;
    const gfloat *row[2][2] = {{r00, r01}, {r00 + px, r01 + px}};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int i = 0; i < 3; ++i) R[jj][kk][i] = row[jj][kk][i];
    const gfloat *ry = b.o + voff(b, cx, yn, cz);
    const gfloat *rz = b.o + voff(b, cx, cy, zn);
    const gfloat *ry1 = ry + pxy, *rz1 = rz + px;
    YN[0][0] = ry[0];
    YN[0][1] = ry[1];
    YN[1][0] = ry1[0];
    YN[1][1] = ry1[1];
    ZN[0][0] = rz[0];
    ZN[0][1] = rz[1];
    ZN[1][0] = rz1[0];
    ZN[1][1] = rz1[1];
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

// Streaming (nontemporal) vector stores: for output that nothing in the
// launch reads back, so the writes do not displace the voxels the march
// re-reads from L2.
__device__ __forceinline__ void store_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(uint32_t *p, uint32_t v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void store_nt(float4 *p, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 u = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(u, reinterpret_cast<f4 *>(p));
}

// Compile-time loop: f(std::integral_constant<int, J>) for J in [J0, N).
template <int J0, int N, class F>
__device__ __forceinline__ void static_for(const F &f) {
  if constexpr (J0 < N) {
    f(std::integral_constant<int, J0>{});
    static_for<J0 + 1, N>(f);
  }
}

// Value of lane gbase + J of the caller's lane group (groups of G aligned
// lanes), on every lane of the group.  Pairs and quads use a DPP quad
// permutation (one VALU move, no LDS round trip); groups of 8 use
// ds_bpermute.  Every lane of the wave must be active.
template <int G, int J>
__device__ __forceinline__ float group_bcast(float v, int gbase) {
  if constexpr (G == 1) {
    return v;
  } else if constexpr (G == 2 || G == 4) {
    constexpr int ctrl = G == 4 ? J * 0x55 : (J | (J << 2) | ((2 + J) << 4) | ((2 + J) << 6));
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
  } else {
    return __shfl(v, gbase + J);
  }
}

template <int G, int J>
__device__ __forceinline__ int group_bcast_i(int v, int gbase) {
  return __float_as_int(group_bcast<G, J>(__int_as_float(v), gbase));
}

// ---- per-pixel helpers shared by the volume and surface raygens ----

__device__ __forceinline__ float sat(float f) { return fminf(fmaxf(f, 0.f), 1.f); }

// heat-map ramp (deviceCode.cpp:300-320)
__device__ inline float3 heat_map(float t) {
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

// Primary ray of pixel (px,py): generateRay (deviceCode.cpp:94-107) at the
// pixel centre (:365).
__device__ __forceinline__ void pixel_ray(const float *dir00, const float *du, const float *dv,
                                          int px, int py, float d[3]) {
  const float su = (float)px + .5f, sv = (float)py + .5f;
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = (dir00[k] + su * du[k]) + sv * dv[k];
  const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d[k] = d[k] / len;
    if (fabsf(d[k]) < 1e-5f) d[k] = 1e-5f;
  }
}

}  // namespace dev
}  // namespace aloha
