// composite.hip -- sort-last compositing of the virtual ranks' images on the
// device, replacing IceT (waikikiRenderer.cpp:894-917, :1043-1079):
//   over        ordered BLEND of premultiplied float RGBA, front to back in
//               the kd visibility order, onto transparent black (dense, or
//               sparse: reading each rank only inside its pixel rectangle);
//   over_rgba8  the same over the ranks' 8-bit images (ubyte colour format);
//   zmin        Z_BUFFER: the smallest depth wins, ties to the lower rank.
// Each is one streaming pass: every rank image is read once with 16-byte
// (float4) or 4-byte coalesced loads and the result written once, so these
// kernels are HBM-bound; each thread owns one pixel.
#include <hip/hip_runtime.h>

#include "device_types.h"

namespace aloha {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t pack8(float f) {
  return (uint32_t)fminf(fmaxf(f * 256.f, 0.f), 255.f);
}
__device__ __forceinline__ uint32_t pack_rgba(float4 c) {
  return pack8(c.x) | (pack8(c.y) << 8) | (pack8(c.z) << 16) | (pack8(c.w) << 24);
}

__global__ __launch_bounds__(kBlock) void over_kernel(const float4 *__restrict__ imgs,
                                                      const CompositeArgs c,
                                                      float4 *__restrict__ out,
                                                      uint32_t *__restrict__ out8) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= c.npix) return;
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int k = 0; k < c.P; ++k) {
    const float4 s = imgs[(int64_t)c.order[k] * c.npix + i];
    const float w = 1.f - C.w;
    C.x = C.x + w * s.x;
    C.y = C.y + w * s.y;
    C.z = C.z + w * s.z;
    C.w = C.w + w * s.w;
  }
  if (out) out[i] = C;
  if (out8) out8[i] = pack_rgba(C);
}

// Sparse 'over' (IceT composites sparse images: runs of empty pixels are
// skipped; here the empty region of rank s is everything outside its pixel
// rectangle c.rect[s]).  Blending a transparent-black pixel leaves C exactly
// unchanged (C + w * 0 = C, w >= 0), so skipping those reads gives the dense
// result bit for bit while reading only the pixels inside the rectangles.
// Grid: the output box [x0, x1) x [y0, y0 + gridDim.y): x covers a row in
// runs of kSparsePix pixels (each thread two pixels kBlock apart), y = row.
// The output goes out with streaming stores.
constexpr int kSparsePix = 2 * kBlock;

__device__ __forceinline__ void store_nt(float4 *p, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 u = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(u, reinterpret_cast<f4 *>(p));
}

// Compact images: each thread issues the loads of up to 8 ranks for both of
// its pixels before it blends them (front to back, as for full-size
// images), so the loads are in flight together instead of one rank's
// latency after the other's (measured on 8 ranks at 1024^2: 5.8 against
// 7.1 us).  Full-size images (W x H x 16 bytes apart) go pixel by pixel,
// rank by rank: loading them in batches measured slower (10.7 against 8.2
// us for back-to-back composites of the same images).
template <bool kCompact>
__global__ __launch_bounds__(kBlock) void over_sparse_kernel(const float4 *__restrict__ imgs,
                                                             const CompositeArgs c, int W,
                                                             int x0, int y0, int x1,
                                                             float4 *__restrict__ out,
                                                             uint32_t *__restrict__ out8) {
  constexpr int kH = kSparsePix / kBlock;  // pixels per thread
  constexpr int kRankBatch = 8;
  const int py = y0 + (int)blockIdx.y;
  if constexpr (!kCompact) {
#pragma unroll
    for (int h = 0; h < kH; ++h) {
      const int px = x0 + (int)(blockIdx.x * kSparsePix + h * kBlock + threadIdx.x);
      if (px >= x1) return;
      const int64_t i = px + (int64_t)W * py;
      float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int k = 0; k < c.P; ++k) {
        const int s = c.order[k];  // k is uniform: scalar loads
        const int32_t *rc = c.rect[s];
        if (px < rc[0] || px >= rc[2] || py < rc[1] || py >= rc[3]) continue;
        const float4 v = imgs[(int64_t)s * c.npix + i];
        const float w = 1.f - C.w;
        C.x = C.x + w * v.x;
        C.y = C.y + w * v.y;
        C.z = C.z + w * v.z;
        C.w = C.w + w * v.w;
      }
      if (out) store_nt(out + i, C);
      if (out8) __builtin_nontemporal_store(pack_rgba(C), out8 + i);
    }
    return;
  }
  int px[kH];
  bool live[kH];
  float4 C[kH];
#pragma unroll
  for (int h = 0; h < kH; ++h) {
    px[h] = x0 + (int)(blockIdx.x * kSparsePix + h * kBlock + threadIdx.x);
    live[h] = px[h] < x1;
    C[h] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  for (int k0 = 0; k0 < c.P; k0 += kRankBatch) {
    float4 v[kH][kRankBatch];
    bool in[kH][kRankBatch];
#pragma unroll
    for (int j = 0; j < kRankBatch; ++j) {
      const bool rank = k0 + j < c.P;  // uniform: scalar loads of the rank's table rows
      const int s = rank ? c.order[k0 + j] : 0;
      const int32_t *rc = c.rect[s];
#pragma unroll
      for (int h = 0; h < kH; ++h) {
        in[h][j] = rank && live[h] && px[h] >= rc[0] && px[h] < rc[2] && py >= rc[1] &&
                   py < rc[3];
        v[h][j] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (in[h][j]) {
          v[h][j] = imgs[c.cbase[s] + (int64_t)(py - rc[1]) * (rc[2] - rc[0]) + (px[h] - rc[0])];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kRankBatch; ++j)
#pragma unroll
      for (int h = 0; h < kH; ++h)
        if (in[h][j]) {  // outside its rectangle a rank is transparent black: C unchanged
          const float w = 1.f - C[h].w;
          C[h].x = C[h].x + w * v[h][j].x;
          C[h].y = C[h].y + w * v[h][j].y;
          C[h].z = C[h].z + w * v[h][j].z;
          C[h].w = C[h].w + w * v[h][j].w;
        }
  }
#pragma unroll
  for (int h = 0; h < kH; ++h) {
    if (!live[h]) continue;
    const int64_t i = px[h] + (int64_t)W * py;
    if (out) store_nt(out + i, C[h]);
    if (out8) __builtin_nontemporal_store(pack_rgba(C[h]), out8 + i);
  }
}

__global__ __launch_bounds__(kBlock) void over_rgba8_kernel(const uint32_t *__restrict__ imgs,
                                                            const CompositeArgs c,
                                                            uint32_t *__restrict__ out8) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= c.npix) return;
  float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int k = 0; k < c.P; ++k) {
    const uint32_t u = imgs[(int64_t)c.order[k] * c.npix + i];
    const float sx = (float)(u & 255u) / 255.f, sy = (float)((u >> 8) & 255u) / 255.f;
    const float sz = (float)((u >> 16) & 255u) / 255.f, sw = (float)(u >> 24) / 255.f;
    const float w = 1.f - C.w;
    C.x = C.x + w * sx;
    C.y = C.y + w * sy;
    C.z = C.z + w * sz;
    C.w = C.w + w * sw;
  }
  out8[i] = pack_rgba(C);
}

__global__ __launch_bounds__(kBlock) void zmin_kernel(const uint32_t *__restrict__ imgs,
                                                      const float *__restrict__ depth, int P,
                                                      int64_t npix, uint32_t *__restrict__ out8,
                                                      float *__restrict__ outz) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= npix) return;
  int best = 0;
  float bz = depth[i];
  for (int r = 1; r < P; ++r) {
    const float z = depth[(int64_t)r * npix + i];
    if (z < bz) {
      bz = z;
      best = r;
    }
  }
  if (out8) out8[i] = imgs[(int64_t)best * npix + i];
  if (outz) outz[i] = bz;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_composite_over(const float4 *imgs, const CompositeArgs &c, float4 *out,
                                 uint32_t *out8, hipStream_t s) {
  over_kernel<<<blocks_for(c.npix), kBlock, 0, s>>>(imgs, c, out, out8);
  return hipGetLastError();
}

hipError_t launch_composite_over_sparse(const float4 *imgs, const CompositeArgs &c, int32_t W,
                                        int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                                        float4 *out, uint32_t *out8, hipStream_t s) {
  if (x0 >= x1 || y0 >= y1) return hipSuccess;
  const dim3 grid((unsigned)((x1 - x0 + kSparsePix - 1) / kSparsePix), (unsigned)(y1 - y0));
  over_sparse_kernel<false><<<grid, kBlock, 0, s>>>(imgs, c, W, x0, y0, x1, out, out8);
  return hipGetLastError();
}

hipError_t launch_composite_over_compact(const float4 *imgs, const CompositeArgs &c, int32_t W,
                                         int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                                         float4 *out, uint32_t *out8, hipStream_t s) {
  if (x0 >= x1 || y0 >= y1) return hipSuccess;
  const dim3 grid((unsigned)((x1 - x0 + kSparsePix - 1) / kSparsePix), (unsigned)(y1 - y0));
  over_sparse_kernel<true><<<grid, kBlock, 0, s>>>(imgs, c, W, x0, y0, x1, out, out8);
  return hipGetLastError();
}

hipError_t launch_composite_over_rgba8(const uint32_t *imgs, const CompositeArgs &c,
                                       uint32_t *out8, hipStream_t s) {
  over_rgba8_kernel<<<blocks_for(c.npix), kBlock, 0, s>>>(imgs, c, out8);
  return hipGetLastError();
}

hipError_t launch_composite_zmin(const uint32_t *imgs, const float *depth, int32_t P,
                                 int64_t npix, uint32_t *out8, float *outz, hipStream_t s) {
  zmin_kernel<<<blocks_for(npix), kBlock, 0, s>>>(imgs, depth, P, npix, out8, outz);
  return hipGetLastError();
}

}  // namespace aloha
