// device_types.h -- POD records shared by the host launchers and the HIP
// kernels of the render-and-composite path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aloha_amd.h"

namespace aloha {

// Brick storage in HBM: f32 voxels, or u8 voxels kept as bytes (decoded at
// sample time through a 256-entry table of u/255.f, the floats Volume::get
// makes), with a replicated border of 1 voxel below and 3 above on every
// axis, so that every index a trilinear tap or the 3-wide gradient footprint
// can form (-1 .. n+2) reads the clamp-to-edge value without a per-load
// clamp.  `origin` points at padded voxel (-1,-1,-1) (the allocation start),
// so tap offsets are non-negative 24-bit products.
constexpr int kPadLo = 1;
constexpr int kPadHi = 3;
// A u8 brick's 3-voxel rows are read 4 bytes at a time: the last row of the
// last slice reads one byte past the padded grid, so its allocation carries
// this many bytes more.
constexpr int kU8Slack = 16;

struct DevBrick {
  float space_lo[3], space_hi[3];  // spaceRange (ray clipping)
  float vox_lo[3];                 // (float)voxelRange.lower (texel origin)
  int32_t n[3];                    // voxel extent voxelRange.size()
  int32_t px;                      // padded x pitch, elements
  int32_t pxy;                     // padded slice pitch, elements
  int32_t dtype;                   // storage encoding: ALOHA_F32 or ALOHA_U8
  int32_t touched_word;            // footprint pass: first 32-bit word of this brick's bitset
  const void *origin;              // float or uint8_t voxels, by dtype
};

// Kernel arguments of one render launch (all virtual ranks at once).
struct MarchArgs {
  float org[3], dir00[3], du[3], dv[3];
  float dt, max_depth, jitter_value, heat_scale;
  float inv_frames, inv_spp;  // 1 / (accum_id + 1), 1 / spp (host IEEE divisions)
  int32_t W, H, spp, accum_id;
  int32_t island_id, num_islands, jitter_mode, flags;
  int32_t P, tf_n;
  uint32_t heat_clock;  // clock delta of every pixel under ALOHA_FLAG_HEAT_CLOCK
  int32_t u8_lut;       // nonzero: a brick of the launch stores u8 voxels (the decode table is staged)
  const DevBrick *bricks;
  const float4 *tf;
  float4 *rgba;
  uint32_t *rgba8;
  float *depth;
  float4 *accum;
  uint32_t *nsamples;
  uint32_t *touched;  // footprint pass only
  int32_t rank_off[ALOHA_MAX_RANKS + 1];
  int32_t rank_bricks[ALOHA_MAX_RANK_BRICKS];
  // per rank slot: pixel rectangle [x0,x1) x [y0,y1) outside of which no
  // primary ray can reach one of the slot's bricks (host-projected, with a
  // margin); pixels outside skip the march and keep the miss result
  int32_t rect[ALOHA_MAX_RANKS][4];
  int32_t order[ALOHA_MAX_RANKS];  // fused composite: rank slots front to back
};

// Persistent per-rank launch: the pixels of the ranks' rectangles in runs of
// up to kRankRun consecutive pixels of one rectangle row, numbered rank after
// rank, row-major (run_off[s] is the first run of rank slot s), the run list
// cut into kRankParts contiguous parts.  Part x belongs to the waves
// w = x (mod kRankParts), which take its runs round-robin.
constexpr int kRankParts = 16;
constexpr int kRankRun = 16;
struct RanksArgs {
  int32_t run_off[ALOHA_MAX_RANKS + 1];
  // compact != 0: the rank images hold only the pixels of each slot's
  // rectangle, slot s's row y at element cbase[s] + (y - y0) * (x1 - x0)
  // (x from x0), and nothing is written outside the rectangles
  int32_t compact;
  int32_t cbase[ALOHA_MAX_RANKS];
  // A frame's composite outputs (W x H; may be NULL): the launch writes them
  // transparent outside the box [box.x, box.z) x [box.y, box.w), which the
  // composite then covers.
  float4 *frame;
  uint32_t *frame8;
  int4 box;
};

// Fused frame tiles (single fused frame and every batched view): 16x4 pixels.
constexpr int kFusedTileW = 16, kFusedTileH = 4;

// Batched form: V views of one rank table in one launch.
// Workgroups are dealt to the GPU's XCDs (each with its own L2) in turn:
// workgroup i runs on XCD i % kXcds.
constexpr int kXcds = 8;

struct BatchArgs {
  MarchArgs m;              // ranks, bricks, TF, params; m.rgba/rgba8/nsamples are V x H x W
  const float *cams;        // V x 12 (org, dir_00, dir_du, dir_dv)
  const int32_t *orders;    // V x P rank slots, front to back
  const int32_t *rects;     // V x P x 4 pixel rectangles
  const int32_t *boxes;     // V x 4: tile bounding box of the view's rectangles (x0, y0, nx, ny)
  const int32_t *box_off;   // V + 1: prefix sums of nx * ny (active work items)
  // persistent schedule: each view's work items cut into kXcds contiguous
  // parts; xoff[x * (V + 1) + v] = the items of parts x of views < v, and
  // counter[x] the next index into the list of parts x
  const int32_t *xoff;
  uint32_t *counter;        // kXcds counters
  int32_t V, total_active;  // views, active work items (box_off[V])
  int32_t fill_blocks;      // non-persistent schedule: blocks that write the rows outside the boxes
  int32_t tile_wl;          // tiles of 64 pixels, (1 << tile_wl) x (64 >> tile_wl): 3, 4 or 5
};

// Differentiable form: one rank marching its bricks, seen from V views.
constexpr int kMaxViews = 16;
constexpr int kMaxDiffBricks = 64;
constexpr int kMaxVisits = 64;

struct DiffArgs {
  float cam[kMaxViews][12];  // org, dir_00, dir_du, dir_dv per view
  float dt, jitter_value, alpha_max;
  int32_t jitter_mode, W, H, V, tf_n, nb;
  int32_t brick_ids[kMaxDiffBricks];
  float *dvox[kMaxDiffBricks];  // per brick slot: dLoss/dvoxel grid (unpadded) or null
  const DevBrick *bricks;
  const float4 *tf;
  float4 *rgba;                 // V x H x W
  uint32_t *nsamples;           // V x H x W
  const float4 *dldc;           // backward: dLoss/dRGBA, V x H x W
  float *dtf;                   // backward: dLoss/dTF, tf_n x 4 (accumulated)
  // forward launch: per view the 16x16 pixel blocks [x0, x0 + nx) x
  // [y0, y0 + ny) (fbox = x0, y0, nx, ny) covering the rectangle its bricks
  // project to -- no ray of a pixel outside reaches a brick -- and
  // block_off[v] = the blocks of the views before v
  int32_t fbox[kMaxViews][4];
  int32_t block_off[kMaxViews + 1];
};

// Backward of the differentiable form: the density gradient goes through
// spatial bins.  Bin = one 16^3 block of a brick's voxels, keyed by the
// lowest corner a sample's seven taps can read, floor(x - .5) per axis
// (clamped to [-1, n]); the block of bin (bx, by, bz) spans corners
// [16 b - 1, 16 b + 15), its taps' voxels [16 b - 1, 16 b + 17): an 18^3 tile.
constexpr int kBinEdge = 16;
constexpr int kTileEdge = kBinEdge + 2;
constexpr int kTileVox = kTileEdge * kTileEdge * kTileEdge;
// Records are 3 float4 per sample: (x, y, z, d/d(value tap)),
// (d/d(gradient) x, y, z, value), (d/d(TF colour) r, g, b, d/d(opacity)).
constexpr int kRecWords = 3;
constexpr int kLdsTfChunk = 1024;  // TF entries whose gradient a bin sums in LDS at once
constexpr int kTfGroupBins = 32;   // bins per group of the TF-gradient reduction
constexpr int kSlotCounters = 8;  // counters the waves take their slots from
constexpr int kLiveShift = 40;    // slot counter bits below its live-wave count
// Sample slots, interleaved by wave: the 64 rays of a wave's 8x8 pixel tile
// share 64 x (their longest sample count) slots from wave_off[w]; backward
// step i of the ray in tile place l (step 0 = its last sample) is slot
// wave_off[w] + 64 i + l, so the 64 lanes of the reverse walk read and write
// one contiguous run of slots per step.  A sample's resampled value and
// gradient, its record and its run mark share its slot.  A run is a maximal
// set of consecutive samples of a ray in one bin; its first sample's slot
// (the head) holds the run's bin in run_bin (-1 elsewhere) and (length,
// density bound, TF bound, rank in the bin) in run_info; its other samples
// sit 64, 128, ... slots below the head.
struct DiffBins {
  int32_t base[kMaxDiffBricks + 1];  // first bin of each brick slot
  int32_t nbx[kMaxDiffBricks], nby[kMaxDiffBricks], nbz[kMaxDiffBricks];
  int32_t nbins;
  uint32_t *count;     // samples per bin (summed over its runs)
  uint32_t *maxd;      // per bin: bits of the largest record density bound (>= 0 or NaN)
  uint32_t *maxt;      // per bin: bits of the largest record TF bound
  uint32_t *off;       // exclusive prefix of count, nbins + 1 entries
  uint32_t total;      // slots in use
  uint32_t cap;        // slots allocated
  int32_t nwaves;      // V x ceil(H / 16) x ceil(W / 16) x 4
  unsigned long long *wave_off;  // per wave: counter << 56 | first slot in the counter's region
  // kSlotCounters counters, each: the slots taken from it (low kLiveShift
  // bits) and the waves with samples among them (high bits)
  unsigned long long *slot_ctr;
  // the waves with samples: counter c's at live[c * live_cap ..], in the
  // order their atomics took; wave i of all of them is list c's entry
  // i - live_pre[c] for live_pre[c] <= i < live_pre[c + 1]
  int32_t *live;
  uint32_t live_cap;
  uint32_t live_pre[kSlotCounters + 1];
  unsigned long long region[kSlotCounters];  // first slot of each counter's region
  float4 *samples;     // per slot: (value, gx, gy, gz)
  float4 *records;     // per slot: kRecWords float4
  int32_t *run_bin;    // per slot
  float4 *run_info;    // per head slot: (length bits, bound d, bound t, rank bits)
  uint32_t *rec_list;  // per bin (at off[bin]): the slots of its samples
  float *tiles;        // nbins x kTileVox density-gradient partial sums
  double *tf_part;     // nbins x 4 tf_n TF-gradient partial sums
};

struct CompositeArgs {
  int32_t P;
  int64_t npix;
  int32_t order[ALOHA_MAX_RANKS];
  // sparse form: rank s's image is transparent outside [x0,x1) x [y0,y1)
  int32_t rect[ALOHA_MAX_RANKS][4];
  // compact sparse images (RanksArgs::compact): slot s's rectangle row y at
  // element cbase[s] + (y - y0) * (x1 - x0)
  int32_t cbase[ALOHA_MAX_RANKS];
};

// Surface path: one rank's triangle geometry (4-wide BVH nodes of 8 float4,
// see TriNode4; triangles of 3 float4 in leaf order, w = primitive id /
// colour id bits) and the launch arguments of all ranks.
constexpr int kTriStack = 96;  // 3 entries per level of the deepest (32) leaf
struct DevTris {
  const float4 *nodes;
  const float4 *tris;
  int32_t root, pad;
};

struct TrisArgs {
  float org[3], dir00[3], du[3], dv[3];
  float max_depth, heat_scale;
  int32_t W, H, spp, accum_id;
  int32_t island_id, num_islands, flags, P;
  uint32_t heat_clock, pad0;  // clock delta of every pixel under ALOHA_FLAG_HEAT_CLOCK
  float4 *rgba;
  uint32_t *rgba8;
  float *depth;
  float4 *accum;
  uint32_t *stats;  // 2 per pixel: nodes visited, triangles tested
  DevTris geo[ALOHA_MAX_RANKS];
  int32_t rect[ALOHA_MAX_RANKS][4];  // per rank: the pixels its geometry can cover
};

// launchers (march.hip, composite.hip)
// group / lanes: lanes per ray (1, 2, 4 or 8) of the march kernels
// footprint pass (voxel bitset of one frame, march_kernel<true>)
hipError_t launch_march_footprint(const MarchArgs &a, hipStream_t s);
// persistent per-rank launch: `waves` one-wave blocks
hipError_t launch_march_ranks(const MarchArgs &a, const RanksArgs &q, int group, int waves,
                              hipStream_t s);
// Dynamic LDS of the march launches: the TF with its zero borders
// (tf_n + 4 float4), a 64-byte brick record per rank_bricks entry,
// rank_off[0..P], the P pixel rectangles and, when a brick of the launch
// stores u8 voxels, the 256-entry decode table.
constexpr size_t kLdsBrickBytes = 64;
constexpr int kU8Lut = 256;
__host__ __device__ inline size_t march_tables_bytes(const MarchArgs &a) {
  return sizeof(float4) * (size_t)(a.tf_n + 4) + kLdsBrickBytes * (size_t)a.rank_off[a.P] +
         sizeof(int32_t) * (size_t)(a.P + 1) + sizeof(int32_t) * 4 * (size_t)a.P +
         (a.u8_lut ? sizeof(float) * kU8Lut : 0);
}
// The persistent per-rank launch adds a list of 64 ray records (48 B each)
// after the tables, at the next 16-byte boundary past them, then the run
// offsets run_off[0..P] and the compact image bases cbase[0..P).
constexpr size_t kRayRecBytes = 48;
__host__ __device__ inline size_t ranks_lds_bytes(const MarchArgs &a) {
  return (march_tables_bytes(a) / 16 + 1) * 16 + 64 * kRayRecBytes +
         sizeof(int32_t) * (size_t)(2 * a.P + 1);
}
// resident one-wave blocks per CU of the persistent per-rank kernel
hipError_t ranks_occupancy(int group, bool u8, size_t lds, int *blocks_per_cu);
hipError_t launch_march_composite(const MarchArgs &a, int group, hipStream_t s);
// lanes: lanes per ray (1, 2, 4, 8); persistent_blocks 0: one block per active tile
hipError_t launch_batch(const BatchArgs &b, int lanes, int persistent_blocks, hipStream_t s);
// decode (to f32 storage) or copy (u8 to u8 storage) and pad one brick
hipError_t launch_pad_brick(const void *src, int dtype, const int32_t n[3], void *dst_base,
                            int store_dtype, int32_t px, int32_t py, int32_t pz, hipStream_t s);
hipError_t launch_diff(const DiffArgs &a, bool backward, hipStream_t s);
// backward in stages: count samples per bin and scan, then the march that
// writes records and the per-bin bounds, then bins into tiles and tiles into
// the gradient grids
hipError_t launch_diff_rays(const DiffArgs &a, const DiffBins &b, hipStream_t s);
hipError_t launch_diff_backward(const DiffArgs &a, const DiffBins &b, hipStream_t s);
// max_dims: the largest extent per axis over the bricks with a gradient grid
hipError_t launch_diff_bins(const DiffArgs &a, const DiffBins &b, const int max_dims[3],
                            hipStream_t s);
hipError_t launch_composite_over(const float4 *imgs, const CompositeArgs &c, float4 *out,
                                 uint32_t *out8, hipStream_t s);
// sparse form: reads rank pixels only inside c.rect (W x H images); writes
// the output pixels of the box [x0, x1) x [y0, y1) only
hipError_t launch_composite_over_sparse(const float4 *imgs, const CompositeArgs &c, int32_t W,
                                        int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                                        float4 *out, uint32_t *out8, hipStream_t s);
// the same over compact rank images (c.cbase, RanksArgs::compact)
hipError_t launch_composite_over_compact(const float4 *imgs, const CompositeArgs &c, int32_t W,
                                         int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                                         float4 *out, uint32_t *out8, hipStream_t s);

hipError_t launch_composite_over_rgba8(const uint32_t *imgs, const CompositeArgs &c,
                                       uint32_t *out8, hipStream_t s);
hipError_t launch_composite_zmin(const uint32_t *imgs, const float *depth, int32_t P,
                                 int64_t npix, uint32_t *out8, float *outz, hipStream_t s);
// tris.hip: per-rank images (composite = false) or the fused depth composite
hipError_t launch_tris(const TrisArgs &a, bool composite, hipStream_t s);

}  // namespace aloha
