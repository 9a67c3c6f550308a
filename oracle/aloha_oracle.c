/*
 * aloha_oracle.c -- CPU restatement of the reference's per-brick
 * render-and-composite path (vtvamr/aloha), used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path
 * (aloha_amd/) never links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   orc_split_volume       chopSuey.cpp:245-323      (VolumeSplitter)
 *   orc_volume_get         volume.h:46-98            (Volume::get value conversion)
 *   orc_partition_kd       waikikiRenderer.cpp:150-270 (Partitioner::partitionKD)
 *   orc_partition_rr       waikikiRenderer.cpp:137-148 (partitionRoundRobin)
 *   orc_composite_order    waikikiRenderer.cpp:281-316 (computeCompositeOrder)
 *   orc_render_rank        deviceCode.cpp:340-416 (renderFrame) with
 *                          :94-107 (generateRay), :199-239 (VolBrickIsect),
 *                          :41-57 (intersect), :245-254 (firstSampleT),
 *                          :256-298 (integrateBrick)
 *   orc_composite_over     waikikiRenderer.cpp:911-916,1043-1077 (IceT ordered
 *                          BLEND of premultiplied colour, front to back)
 *   orc_composite_zmin     waikikiRenderer.cpp:903-909 (IceT Z_BUFFER)
 *   orc_mesh_*             deviceCode.cpp:59-67,176-188,366-375 (ModelCH,
 *                          randomColor, headlight shading) by brute force
 *
 * Conventions this build fixes where the reference's behaviour lives in
 * un-vendored code (OWL, fakeOwl, CUDA texture units, IceT):
 *   - texture addressing: texel space x = pos - voxelRange.lower, clamp to
 *     [0,n-1] per axis of the brick, trilinear with lerp(a,b,w)=fma(w,b-a,a),
 *     x first, then y, then z;
 *   - TF: x = v*N - 0.5, border entries T[-1] = T[N] = 0, same lerp;
 *   - normalize(v) = v / sqrt(dot(v,v)); dot = (x*x + y*y) + z*z;
 *   - jitter RNG: OWL's LCG<4> -- 4 TEA rounds over (pixelID, accumID)
 *     seed the Numerical-Recipes LCG, draws (state & 0xFFFFFF)/2^24 (the
 *     published construction; OWL's source is not in the reference);
 *   - 8-bit pack: min(255, max(0, int(f*256))), R in the low byte;
 *   - composite: premultiplied 'over' onto transparent black.
 * Compile with -ffp-contract=off: every float operation below is rounded
 * exactly as written, which the HIP kernels reproduce.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_U8 1
#define ORC_U16 2
#define ORC_F32 4

typedef struct {
  int32_t cell_lo[3], cell_hi[3];
  int32_t vox_lo[3], vox_hi[3];
  float space_lo[3], space_hi[3];
  int32_t dtype;
  int32_t pad_;
  const void *voxels; /* x-fastest, dims = vox_hi - vox_lo */
} orc_brick;

typedef struct {
  float org[3], dir00[3], du[3], dv[3];
  float dt, max_depth;
  int32_t W, H;
  int32_t spp, accum_id;
  int32_t island_id, num_islands;
  int32_t jitter_mode; /* 0: per-pixel RNG, 1: fixed value */
  float jitter_value;
  int32_t crosshair; /* 1: invert row H/2 and column W/2 in the 8-bit output */
  int32_t tf_n;
  const float *tf; /* tf_n x RGBA */
  int32_t heat;         /* 1: blend the heat map of every pixel's clock delta heat_clock */
  float heat_scale;     /* heatMapScale */
  uint32_t heat_clock;  /* the clock delta (clock_end - clock_begin) of every pixel */
  int32_t pad_;
} orc_params;

typedef struct {
  int32_t axis;
  float plane;
  int32_t child1, child2;
} orc_kdnode;

/* ------------------------------------------------------------------ */
/* RNG: owl::common::LCG<4> (deviceCode.cpp:12, seeded per pixel at :354, */
/* one draw per brick visit at :386).  OWL is not vendored; its LCG<N> is  */
/* the published construction of Zafar, Olano & Curtis, "GPU random       */
/* numbers via the tiny encryption algorithm" (HPG 2010), as the OptiX SDK */
/* samples ship it (random.h, tea<N> + lcg + rnd):                         */
/*   seed  = v0 after N rounds of TEA (Wheeler & Needham 1994) over        */
/*           (v0, v1) = (pixelID, accumID) with the key                    */
/*           {0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e};             */
/*   draw  = state = 1664525*state + 1013904223 (Numerical Recipes),       */
/*           (state & 0xFFFFFF) / 2^24.                                    */
/* ------------------------------------------------------------------ */
/* TEA encryption of (v[0], v[1]) under key k, `rounds` cycles. */
void orc_tea(uint32_t v[2], const uint32_t k[4], int32_t rounds) {
  uint32_t v0 = v[0], v1 = v[1], sum = 0;
  for (int32_t i = 0; i < rounds; ++i) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + k[0]) ^ (v1 + sum) ^ ((v1 >> 5) + k[1]);
    v1 += ((v0 << 4) + k[2]) ^ (v0 + sum) ^ ((v0 >> 5) + k[3]);
  }
  v[0] = v0;
  v[1] = v1;
}
static const uint32_t kLcgKey[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
static inline uint32_t rng_seed(uint32_t pixel, uint32_t frame) {
  uint32_t v[2] = {pixel, frame};
  orc_tea(v, kLcgKey, 4);
  return v[0];
}
static inline float rng_next(uint32_t *s) {
  *s = *s * 1664525u + 1013904223u;
  return (float)(*s & 0x00FFFFFFu) / (float)0x01000000;
}
uint32_t orc_rng_seed(uint32_t pixel, uint32_t frame) { return rng_seed(pixel, frame); }
float orc_rng_draw(uint32_t *s) { return rng_next(s); }

/* ------------------------------------------------------------------ */
/* voxel decode (volume.h:76-86)                                       */
/* ------------------------------------------------------------------ */
static inline float decode(const orc_brick *b, size_t idx) {
  if (b->dtype == ORC_U8) return (float)((const uint8_t *)b->voxels)[idx] / 255.f;
  if (b->dtype == ORC_U16) return (float)((const uint16_t *)b->voxels)[idx] / 65535.f;
  return ((const float *)b->voxels)[idx];
}

/* Volume::get restated: read the sub-box [first,last) of an x-fastest raw
 * volume of `dims` with bytes-per-voxel bpc, convert to float, and extend
 * the value range (volume.h:46-98). */
int orc_volume_get(const void *raw, const int32_t dims[3], int bpc, const int32_t first[3],
                   const int32_t last[3], float *out, float value_range[2]) {
  if (bpc != 1 && bpc != 2 && bpc != 4) return -1;
  float lo = FLT_MAX, hi = -FLT_MAX;
  size_t off = 0;
  for (int z = first[2]; z != last[2]; ++z)
    for (int y = first[1]; y != last[1]; ++y)
      for (int x = first[0]; x != last[0]; ++x) {
        size_t i = (size_t)z * dims[0] * (size_t)dims[1] + (size_t)y * dims[0] + x;
        float v;
        if (bpc == 1) v = (float)((const uint8_t *)raw)[i] / 255.f;
        else if (bpc == 2) v = (float)((const uint16_t *)raw)[i] / 65535.f;
        else v = ((const float *)raw)[i];
        out[off++] = v;
        if (v < lo) lo = v;
        if (v > hi) hi = v;
      }
  if (value_range) {
    if (lo < value_range[0]) value_range[0] = lo;
    if (hi > value_range[1]) value_range[1] = hi;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* bricking: VolumeSplitter (chopSuey.cpp:245-323)                     */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t cell_lo[3], cell_hi[3], vox_lo[3], vox_hi[3];
  float space_lo[3], space_hi[3];
} orc_domain;

static float fvolume(const float lo[3], const float hi[3]) {
  /* owl box3f::volume(): size.x*size.y*size.z */
  float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
  return sx * sy * sz;
}
static long long ivolume(const int32_t lo[3], const int32_t hi[3]) {
  return (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
}

/* Returns the number of bricks written to out (<= cap), or -1. */
int orc_split_volume(const int32_t dims[3], int n, orc_domain *out, int cap) {
  orc_domain *cl = (orc_domain *)malloc(sizeof(orc_domain) * (size_t)(cap + 2));
  int count = 0;
  orc_domain d;
  memset(&d, 0, sizeof d);
  for (int a = 0; a < 3; ++a) {
    d.cell_hi[a] = dims[a];
    d.vox_hi[a] = dims[a] + 1;
    d.space_hi[a] = (float)dims[a];
  }
  cl[count++] = d;
  do {
    /* pick the brick with the largest spaceRange volume, strict >, first wins */
    int pick = 0;
    float maxv = 0.f;
    for (int i = 0; i < count; ++i) {
      float v = fvolume(cl[i].space_lo, cl[i].space_hi);
      if (v > maxv) { pick = i; maxv = v; }
    }
    orc_domain dom = cl[pick];
    memmove(cl + pick, cl + pick + 1, sizeof(orc_domain) * (size_t)(count - pick - 1));
    --count;
    int sz[3] = {dom.cell_hi[0] - dom.cell_lo[0], dom.cell_hi[1] - dom.cell_lo[1],
                 dom.cell_hi[2] - dom.cell_lo[2]};
    int axis = 0;
    if (sz[1] > sz[0] && sz[1] >= sz[2]) axis = 1;
    else if (sz[2] > sz[0] && sz[2] >= sz[1]) axis = 2;
    /* int splitPlane = lower + size*.5f (float product, truncating conversion) */
    int plane = (int)((float)dom.cell_lo[axis] + (float)sz[axis] * .5f);
    orc_domain L = dom, R = dom;
    L.cell_hi[axis] = plane;
    R.cell_lo[axis] = plane;
    orc_domain *lr[2] = {&L, &R};
    for (int k = 0; k < 2; ++k) {
      orc_domain *c = lr[k];
      for (int a = 0; a < 3; ++a) {
        c->vox_lo[a] = c->cell_lo[a];
        int up = c->cell_hi[a] + 1;
        c->vox_hi[a] = up < dims[a] ? up : dims[a];
        c->space_lo[a] = (float)c->cell_lo[a];
        c->space_hi[a] = (float)c->cell_hi[a];
      }
    }
    int before = count;
    if (ivolume(L.cell_lo, L.cell_hi) > 0) cl[count++] = L;
    if (ivolume(R.cell_lo, R.cell_hi) > 0) cl[count++] = R;
    if (count > cap || count < before + 2) { free(cl); return -1; }
  } while (count < n);
  memcpy(out, cl, sizeof(orc_domain) * (size_t)count);
  free(cl);
  return count;
}

/* ------------------------------------------------------------------ */
/* Partitioner (waikikiRenderer.cpp:126-338)                           */
/* ------------------------------------------------------------------ */
typedef struct {
  float lo[3], hi[3];
  int *ids;
  int n;
  int kd;
} part_set;

static void box_extend(float lo[3], float hi[3], const float *dlo, const float *dhi) {
  for (int a = 0; a < 3; ++a) {
    if (dlo[a] < lo[a]) lo[a] = dlo[a];
    if (dhi[a] > hi[a]) hi[a] = dhi[a];
  }
}

/* domains: n x 6 floats (lower xyz, upper xyz).  Outputs:
 *   rank_count[P], rank_ids[n] (concatenated per rank, in assignment order),
 *   kd nodes (returns node count, or -1 if kd_cap is too small). */
int orc_partition_kd(const float *domains, int n, int P, int *rank_count, int *rank_ids,
                     orc_kdnode *kd, int kd_cap) {
  int per_rank = n / P;
  part_set *open = (part_set *)calloc((size_t)(2 * n + 4), sizeof(part_set));
  part_set *asg = (part_set *)calloc((size_t)(P + 1), sizeof(part_set));
  int nopen = 0, nasg = 0, nkd = 0;
  part_set all;
  for (int a = 0; a < 3; ++a) { all.lo[a] = 1e30f; all.hi[a] = -1e30f; }
  all.ids = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  all.n = 0;
  all.kd = 0;
  for (int i = 0; i < n; ++i) {
    box_extend(all.lo, all.hi, domains + 6 * i, domains + 6 * i + 3);
    all.ids[all.n++] = i;
  }
  open[nopen++] = all;
  if (nkd >= kd_cap) return -1;
  kd[nkd++] = (orc_kdnode){0, 0.f, INT_MAX, INT_MAX};
  while (nopen > 0) {
    int pick = 0, maxn = 0;
    for (int i = 0; i < nopen; ++i)
      if (open[i].n > maxn) { pick = i; maxn = open[i].n; }
    part_set cs = open[pick];
    memmove(open + pick, open + pick + 1, sizeof(part_set) * (size_t)(nopen - pick - 1));
    --nopen;
    if (cs.n <= per_rank || nasg == P - 1) {
      int rank = nasg;
      kd[cs.kd].child1 = ~rank;
      kd[cs.kd].child2 = ~rank;
      asg[nasg++] = cs;
      if (nasg == P) {
        part_set *last = &asg[nasg - 1];
        int extra = 0;
        for (int i = 0; i < nopen; ++i) extra += open[i].n;
        int *ids = (int *)malloc(sizeof(int) * (size_t)(last->n + extra + 1));
        memcpy(ids, last->ids, sizeof(int) * (size_t)last->n);
        int m = last->n;
        for (int i = 0; i < nopen; ++i) {
          memcpy(ids + m, open[i].ids, sizeof(int) * (size_t)open[i].n);
          m += open[i].n;
          free(open[i].ids);
        }
        free(last->ids);
        last->ids = ids;
        last->n = m;
        nopen = 0;
        break;
      }
      continue;
    }
    float s[3] = {cs.hi[0] - cs.lo[0], cs.hi[1] - cs.lo[1], cs.hi[2] - cs.lo[2]};
    int axis = 0;
    if (s[1] > s[0] && s[1] >= s[2]) axis = 1;
    else if (s[2] > s[0] && s[2] >= s[1]) axis = 2;
    float plane = cs.lo[axis] + s[axis] * .5f;
    part_set L, R;
    for (int a = 0; a < 3; ++a) { L.lo[a] = R.lo[a] = 1e30f; L.hi[a] = R.hi[a] = -1e30f; }
    L.ids = (int *)malloc(sizeof(int) * (size_t)(cs.n + 1));
    R.ids = (int *)malloc(sizeof(int) * (size_t)(cs.n + 1));
    L.n = R.n = 0;
    for (int i = 0; i < cs.n; ++i) {
      int id = cs.ids[i];
      const float *d = domains + 6 * id;
      /* box3f::center() = (lower+upper)*.5f */
      float c = (d[axis] + d[3 + axis]) * .5f;
      if (c < plane) { box_extend(L.lo, L.hi, d, d + 3); L.ids[L.n++] = id; }
      else { box_extend(R.lo, R.hi, d, d + 3); R.ids[R.n++] = id; }
    }
    free(cs.ids);
    /* a split that separates nothing repeats forever in the reference */
    if (L.n == 0 || R.n == 0) { free(L.ids); free(R.ids); return -2; }
    kd[cs.kd].axis = axis;
    kd[cs.kd].plane = plane;
    if (nkd + 2 > kd_cap) return -1;
    L.kd = nkd;
    kd[cs.kd].child1 = nkd;
    kd[nkd++] = (orc_kdnode){0, 0.f, INT_MAX, INT_MAX};
    open[nopen++] = L;
    R.kd = nkd;
    kd[cs.kd].child2 = nkd;
    kd[nkd++] = (orc_kdnode){0, 0.f, INT_MAX, INT_MAX};
    open[nopen++] = R;
  }
  int m = 0;
  for (int r = 0; r < P; ++r) {
    rank_count[r] = r < nasg ? asg[r].n : 0;
    for (int j = 0; r < nasg && j < asg[r].n; ++j) rank_ids[m++] = asg[r].ids[j];
    if (r < nasg) free(asg[r].ids);
  }
  free(open);
  free(asg);
  return nkd;
}

/* partitionRoundRobin: rank = i / ceil(n/P) (a block distribution). */
int orc_partition_rr(int n, int P, int *rank_of) {
  int per = (n + P - 1) / P;
  for (int i = 0; i < n; ++i) rank_of[i] = i / per;
  return 0;
}

/* computeCompositeOrder: depth-first walk, near child first. */
int orc_composite_order(const orc_kdnode *kd, int nkd, const float ref[3], int *order, int cap) {
  if (nkd <= 0) return 0;
  int *stack = (int *)malloc(sizeof(int) * (size_t)(nkd + 2));
  int sp = 0, addr = 0, m = 0;
  stack[sp++] = addr;
  while (sp > 0) {
    orc_kdnode node = kd[addr];
    if (node.child1 < 0 && node.child2 < 0) {
      if (m < cap) order[m] = ~node.child1;
      ++m;
      addr = stack[--sp];
    } else if (node.child1 == INT_MAX) {
      addr = stack[--sp];
    } else {
      if (ref[node.axis] < node.plane) {
        addr = node.child1;
        stack[sp++] = node.child2;
      } else {
        addr = node.child2;
        stack[sp++] = node.child1;
      }
    }
  }
  free(stack);
  return m;
}

/* ------------------------------------------------------------------ */
/* sampling                                                            */
/* ------------------------------------------------------------------ */
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float lerpf_(float a, float b, float w) { return fmaf(w, b - a, a); }
static inline double lerpd_(double a, double b, double w) { return fma(w, b - a, a); }

typedef struct {
  const orc_brick *b;
  int n[3];
  uint8_t *touched; /* optional per-voxel marks */
} bview;

static inline float vox(const bview *v, int x, int y, int z) {
  x = clampi(x, 0, v->n[0] - 1);
  y = clampi(y, 0, v->n[1] - 1);
  z = clampi(z, 0, v->n[2] - 1);
  size_t idx = ((size_t)z * v->n[1] + y) * v->n[0] + x;
  if (v->touched) __atomic_store_n(&v->touched[idx], 1, __ATOMIC_RELAXED);
  return decode(v->b, idx);
}

static inline int floor_to_int(float f) {
  /* the values here are bounded by the brick; guard the conversion only */
  if (!(f > -1e8f)) f = -1e8f;
  if (f > 1e8f) f = 1e8f;
  return (int)f;
}

static float tri_f(const bview *v, float x, float y, float z) {
  float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  float wx = x - fx, wy = y - fy, wz = z - fz;
  int ix = floor_to_int(fx), iy = floor_to_int(fy), iz = floor_to_int(fz);
  float c00 = lerpf_(vox(v, ix, iy, iz), vox(v, ix + 1, iy, iz), wx);
  float c10 = lerpf_(vox(v, ix, iy + 1, iz), vox(v, ix + 1, iy + 1, iz), wx);
  float c01 = lerpf_(vox(v, ix, iy, iz + 1), vox(v, ix + 1, iy, iz + 1), wx);
  float c11 = lerpf_(vox(v, ix, iy + 1, iz + 1), vox(v, ix + 1, iy + 1, iz + 1), wx);
  float c0 = lerpf_(c00, c10, wy);
  float c1 = lerpf_(c01, c11, wy);
  return lerpf_(c0, c1, wz);
}

static double tri_d(const bview *v, double x, double y, double z) {
  double fx = floor(x), fy = floor(y), fz = floor(z);
  double wx = x - fx, wy = y - fy, wz = z - fz;
  int ix = floor_to_int((float)fx), iy = floor_to_int((float)fy), iz = floor_to_int((float)fz);
  bview nv = *v;
  nv.touched = NULL;
  double c00 = lerpd_(vox(&nv, ix, iy, iz), vox(&nv, ix + 1, iy, iz), wx);
  double c10 = lerpd_(vox(&nv, ix, iy + 1, iz), vox(&nv, ix + 1, iy + 1, iz), wx);
  double c01 = lerpd_(vox(&nv, ix, iy, iz + 1), vox(&nv, ix + 1, iy, iz + 1), wx);
  double c11 = lerpd_(vox(&nv, ix, iy + 1, iz + 1), vox(&nv, ix + 1, iy + 1, iz + 1), wx);
  double c0 = lerpd_(c00, c10, wy);
  double c1 = lerpd_(c01, c11, wy);
  return lerpd_(c0, c1, wz);
}

/* TF lookup: 1D linear filter, normalized coords, border colour 0
 * (waikikiRenderer.cpp:993-1005; deviceCode.cpp:289). */
static void tf_f(const orc_params *p, float v, float out[4]) {
  int n = p->tf_n;
  float xt = v * (float)n - 0.5f;
  float fi = floorf(xt);
  float w = xt - fi;
  float fc = fi < -2.f ? -2.f : (fi > (float)n ? (float)n : fi);
  if (!(fi == fi)) fc = -2.f;
  int i = (int)fc;
  for (int c = 0; c < 4; ++c) {
    float e0 = (i >= 0 && i < n) ? p->tf[4 * i + c] : 0.f;
    float e1 = (i + 1 >= 0 && i + 1 < n) ? p->tf[4 * (i + 1) + c] : 0.f;
    out[c] = lerpf_(e0, e1, w);
  }
}
static void tf_d(const orc_params *p, double v, double out[4]) {
  int n = p->tf_n;
  double xt = v * (double)n - 0.5;
  double fi = floor(xt);
  double w = xt - fi;
  double fc = fi < -2. ? -2. : (fi > (double)n ? (double)n : fi);
  if (!(fi == fi)) fc = -2.;
  int i = (int)fc;
  for (int c = 0; c < 4; ++c) {
    double e0 = (i >= 0 && i < n) ? p->tf[4 * i + c] : 0.;
    double e1 = (i + 1 >= 0 && i + 1 < n) ? p->tf[4 * (i + 1) + c] : 0.;
    out[c] = lerpd_(e0, e1, w);
  }
}

/* slab test (deviceCode.cpp:41-57) */
static int slab(const float o[3], const float d[3], float tmin, float tmax, const float lo[3],
                const float hi[3], float *t0, float *t1) {
  float nr[3], fr[3];
  for (int a = 0; a < 3; ++a) {
    float l = (lo[a] - o[a]) / d[a];
    float h = (hi[a] - o[a]) / d[a];
    nr[a] = fminf(l, h);
    fr[a] = fmaxf(l, h);
  }
  *t0 = fmaxf(tmin, fmaxf(fmaxf(nr[0], nr[1]), nr[2]));
  *t1 = fminf(tmax, fminf(fminf(fr[0], fr[1]), fr[2]));
  return *t0 < *t1;
}

typedef struct {
  float c[4];
  double c64[4];
  uint32_t nsamples;
  int want64;
} march_state;

/* integrateBrick (deviceCode.cpp:256-298) */
static void integrate(const orc_params *p, const bview *v, const float o[3], const float d[3],
                      float tmin, float jitter, march_state *st) {
  const orc_brick *b = v->b;
  float t0, t1;
  if (!slab(o, d, tmin, 1e10f, b->space_lo, b->space_hi, &t0, &t1)) return;
  const float dt = p->dt;
  /* firstSampleT (deviceCode.cpp:245-254) */
  float segs = floorf((t0 - dt * jitter) / dt);
  float t = dt * (jitter + segs);
  if (t < t0) t += dt;
  const float vl[3] = {(float)b->vox_lo[0], (float)b->vox_lo[1], (float)b->vox_lo[2]};
  for (; t < t1 && st->c[3] < .99f; t += dt) {
    st->nsamples++;
    float x = (o[0] + t * d[0]) - vl[0];
    float y = (o[1] + t * d[1]) - vl[1];
    float z = (o[2] + t * d[2]) - vl[2];
    float gx = tri_f(v, x + .5f, y, z) - tri_f(v, x - .5f, y, z);
    float gy = tri_f(v, x, y + .5f, z) - tri_f(v, x, y - .5f, z);
    float gz = tri_f(v, x, y, z + .5f) - tri_f(v, x, y, z - .5f);
    float ex = gx + 1e-4f, ey = gy + 1e-4f, ez = gz + 1e-4f;
    float len = sqrtf(ex * ex + ey * ey + ez * ez);
    float nx = gx / len, ny = gy / len, nz = gz / len;
    float val = tri_f(v, x, y, z);
    float s[4];
    tf_f(p, val, s);
    float shade = .1f + .9f * fabsf(nx * d[0] + ny * d[1] + nz * d[2]);
    float w = (1.f - st->c[3]) * s[3];
    st->c[0] = st->c[0] + (w * s[0]) * shade;
    st->c[1] = st->c[1] + (w * s[1]) * shade;
    st->c[2] = st->c[2] + (w * s[2]) * shade;
    st->c[3] = st->c[3] + w;
    if (st->want64) {
      double X = ((double)o[0] + (double)t * (double)d[0]) - (double)vl[0];
      double Y = ((double)o[1] + (double)t * (double)d[1]) - (double)vl[1];
      double Z = ((double)o[2] + (double)t * (double)d[2]) - (double)vl[2];
      double Gx = tri_d(v, X + .5, Y, Z) - tri_d(v, X - .5, Y, Z);
      double Gy = tri_d(v, X, Y + .5, Z) - tri_d(v, X, Y - .5, Z);
      double Gz = tri_d(v, X, Y, Z + .5) - tri_d(v, X, Y, Z - .5);
      double Ex = Gx + 1e-4, Ey = Gy + 1e-4, Ez = Gz + 1e-4;
      double L = sqrt(Ex * Ex + Ey * Ey + Ez * Ez);
      double Val = tri_d(v, X, Y, Z);
      double S[4];
      tf_d(p, Val, S);
      double sh = .1 + .9 * fabs((Gx / L) * d[0] + (Gy / L) * d[1] + (Gz / L) * d[2]);
      double W = (1. - st->c64[3]) * S[3];
      for (int c = 0; c < 3; ++c) st->c64[c] += W * S[c] * sh;
      st->c64[3] += W;
    }
  }
}

/* ---- probes of the building blocks, for closed-form tests ---------- */
/* value tap and central differences at texel-space (x,y,z) of brick b:
 * out = {value, gx, gy, gz}; out64 the same in double (may be NULL) */
void orc_sample_probe(const orc_brick *b, float x, float y, float z, float out[4],
                      double out64[4]) {
  bview v;
  v.b = b;
  for (int a = 0; a < 3; ++a) v.n[a] = b->vox_hi[a] - b->vox_lo[a];
  v.touched = NULL;
  out[0] = tri_f(&v, x, y, z);
  out[1] = tri_f(&v, x + .5f, y, z) - tri_f(&v, x - .5f, y, z);
  out[2] = tri_f(&v, x, y + .5f, z) - tri_f(&v, x, y - .5f, z);
  out[3] = tri_f(&v, x, y, z + .5f) - tri_f(&v, x, y, z - .5f);
  if (out64) {
    out64[0] = tri_d(&v, x, y, z);
    out64[1] = tri_d(&v, x + .5, y, z) - tri_d(&v, x - .5, y, z);
    out64[2] = tri_d(&v, x, y + .5, z) - tri_d(&v, x, y - .5, z);
    out64[3] = tri_d(&v, x, y, z + .5) - tri_d(&v, x, y, z - .5);
  }
}
void orc_tf_probe(const float *tf, int n, float v, float out[4]) {
  orc_params p;
  memset(&p, 0, sizeof p);
  p.tf = tf;
  p.tf_n = n;
  tf_f(&p, v, out);
}
float orc_first_sample_t(float t0, float dt, float jitter) {
  float segs = floorf((t0 - dt * jitter) / dt);
  float t = dt * (jitter + segs);
  if (t < t0) t += dt;
  return t;
}
int orc_slab_probe(const float o[3], const float d[3], float tmin, float tmax, const float lo[3],
                   const float hi[3], float t[2]) {
  return slab(o, d, tmin, tmax, lo, hi, &t[0], &t[1]);
}

/* One pixel of renderFrame for one rank whose bricks are b[0..nb). */
static void render_pixel(const orc_params *p, const bview *bv, int nb, int px, int py,
                         float acc[4], double acc64[4], float *zout, uint32_t *ns, int want64) {
  const int W = p->W;
  const int pixel = px + W * py;
  /* island mask with the reference's gridSize formula (deviceCode.cpp:346-361) */
  /* gridSize.x = iDivUp(pixel.x, W) is 1 for every column but the first */
  const int gx = (px + W - 1) / W;
  const int block = px / 16 + gx * (py / 16);
  uint32_t rng = rng_seed((uint32_t)pixel, (uint32_t)p->accum_id);
  float z = 1.f;
  for (int c = 0; c < 4; ++c) { acc[c] = 0.f; acc64[c] = 0.; }
  *ns = 0;
  if (block % p->num_islands == p->island_id) {
    for (int s = 0; s < p->spp; ++s) {
      /* generateRay (deviceCode.cpp:94-107), pixel centre (:365) */
      const float su = (float)px + .5f, sv = (float)py + .5f;
      float d[3];
      for (int a = 0; a < 3; ++a) d[a] = (p->dir00[a] + su * p->du[a]) + sv * p->dv[a];
      float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int a = 0; a < 3; ++a) d[a] = d[a] / len;
      for (int a = 0; a < 3; ++a)
        if (fabsf(d[a]) < 1e-5f) d[a] = 1e-5f;
      march_state st;
      memset(&st, 0, sizeof st);
      st.want64 = want64;
      float tmin = 0.f;
      int hit;
      do {
        /* closest reported hit among this rank's bricks (VolBrickIsect),
         * ties to the lowest brick index */
        hit = -1;
        float best = 0.f, bnear = 0.f, bfar = 0.f;
        for (int i = 0; i < nb; ++i) {
          float t0, t1, rep, tn;
          if (!slab(p->org, d, tmin, 1e10f, bv[i].b->space_lo, bv[i].b->space_hi, &t0, &t1))
            continue;
          if (t0 > tmin) { rep = t0; tn = t0; }
          else if (t1 > tmin) { rep = t1; tn = 0.f; }
          else continue;
          if (hit < 0 || rep < best) { hit = i; best = rep; bnear = tn; bfar = t1; }
        }
        if (hit >= 0) {
          bnear = bnear + tmin;
          const float j = p->jitter_mode == 1 ? p->jitter_value : rng_next(&rng);
          integrate(p, &bv[hit], p->org, d, tmin, j, &st);
          z = fminf(z, bnear / p->max_depth);
          tmin = bfar + 1e-20f;
        }
      } while (hit >= 0 && st.c[3] < .99f);
      for (int c = 0; c < 4; ++c) { acc[c] = acc[c] + st.c[c]; acc64[c] += st.c64[c]; }
      *ns += st.nsamples;
    }
  }
  *zout = z;
}

/* heatMap = temperature_to_rgb (deviceCode.cpp:300-320) */
static float sat_(float f) { return fminf(fmaxf(f, 0.f), 1.f); }
static void heat_rgb(float t, float out[3]) {
  const float K = 4.0f / 6.0f;
  const float h = K - K * t;
  const float v = .5f + 0.5f * t;
  const float s = sat_(h) * 6.0f; /* hue_to_rgb(h) */
  const float r = sat_(fabsf(s - 3.f) - 1.0f);
  const float g = sat_(2.0f - fabsf(s - 2.0f));
  const float b = sat_(2.0f - fabsf(s - 4.0f));
  out[0] = v * r, out[1] = v * g, out[2] = v * b;
}

/* renderFrame's heat-map blend (deviceCode.cpp:399-403) with the pixel's
 * clock delta given: t = delta * (heatMapScale / spp) (the uint64 delta
 * converted to float), then accum = over((heatMap(t), .5), accum) with
 * over(A, B) = A + (1 - A.w) B (:36-39). */
static void heat_blend(const orc_params *p, float acc[4], double acc64[4]) {
  const float t = (float)(uint64_t)p->heat_clock * (p->heat_scale / (float)p->spp);
  float h[3];
  heat_rgb(t, h);
  const float k = 1.f - .5f;
  for (int c = 0; c < 3; ++c) acc[c] = h[c] + k * acc[c];
  acc[3] = .5f + k * acc[3];
  if (acc64) {
    for (int c = 0; c < 3; ++c) acc64[c] = (double)h[c] + .5 * acc64[c];
    acc64[3] = .5 + .5 * acc64[3];
  }
}

static inline uint32_t pack8(float f) {
  float g = f * 256.f;
  g = fmaxf(g, 0.f);
  g = fminf(g, 255.f);
  return (uint32_t)g;
}
uint32_t orc_make_rgba(const float c[4]) {
  return pack8(c[0]) | (pack8(c[1]) << 8) | (pack8(c[2]) << 16) | (pack8(c[3]) << 24);
}

/* Render one rank.  All outputs W*H (rgba: x4), each may be NULL.
 *   accum  in/out float4 running sum (read when accum_id > 0, always written)
 *   touched[i]: per-voxel marks of brick i (may be NULL)
 * rows [row0,row1) only (others untouched), so a bounded sample can be timed. */
int orc_render_rank(const orc_params *p, const orc_brick *bricks, int nb, int row0, int row1,
                    float *rgba, uint32_t *rgba8, float *depth, float *accum,
                    uint32_t *nsamples, double *rgba64, uint8_t **touched, int nthreads) {
  if (!p || p->W <= 0 || p->H <= 0 || p->spp < 1 || p->num_islands < 1 || p->tf_n < 1)
    return -1;
  bview bv[256];
  if (nb > 256) return -2;
  for (int i = 0; i < nb; ++i) {
    bv[i].b = &bricks[i];
    for (int a = 0; a < 3; ++a) bv[i].n[a] = bricks[i].vox_hi[a] - bricks[i].vox_lo[a];
    bv[i].touched = touched ? touched[i] : NULL;
  }
  if (row0 < 0) row0 = 0;
  if (row1 > p->H) row1 = p->H;
  const int W = p->W, H = p->H;
  const float inv_frames = 1.f / (float)(p->accum_id + 1);
  const float inv_spp = 1.f / (float)p->spp;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int py = row0; py < row1; ++py) {
    for (int px = 0; px < W; ++px) {
      const size_t pix = (size_t)py * W + px;
      float acc[4], z;
      double acc64[4];
      uint32_t ns;
      render_pixel(p, bv, nb, px, py, acc, acc64, &z, &ns, rgba64 != NULL);
      if (depth) depth[pix] = z;
      if (nsamples) nsamples[pix] = ns;
      if (p->heat) heat_blend(p, acc, acc64);
      if (accum) {
        if (p->accum_id > 0)
          for (int c = 0; c < 4; ++c) acc[c] = acc[c] + accum[4 * pix + c];
        for (int c = 0; c < 4; ++c) accum[4 * pix + c] = acc[c];
      }
      float fr[4];
      for (int c = 0; c < 4; ++c) fr[c] = (acc[c] * inv_frames) * inv_spp;
      if (rgba)
        for (int c = 0; c < 4; ++c) rgba[4 * pix + c] = fr[c];
      if (rgba64)
        for (int c = 0; c < 4; ++c)
          rgba64[4 * pix + c] = acc64[c] / (double)(p->accum_id + 1) / (double)p->spp;
      if (rgba8) {
        float cc[4];
        const int cross = p->crosshair && (px == W / 2 || py == H / 2);
        for (int c = 0; c < 4; ++c) {
          float a = acc[c] * inv_frames;
          if (cross) a = 1.f - a;
          cc[c] = a * inv_spp;
        }
        rgba8[pix] = orc_make_rgba(cc);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* composite                                                           */
/* ------------------------------------------------------------------ */
/* Ordered premultiplied 'over', front to back in `order` (P rank ids),
 * onto transparent black.  imgs: P x npix x 4 floats (rank-major). */
int orc_composite_over(const float *imgs, const int *order, int P, long long npix, float *out,
                       uint32_t *out8) {
  for (long long i = 0; i < npix; ++i) {
    float C[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < P; ++k) {
      const float *c = imgs + ((size_t)order[k] * (size_t)npix + (size_t)i) * 4;
      const float w = 1.f - C[3];
      C[0] = C[0] + w * c[0];
      C[1] = C[1] + w * c[1];
      C[2] = C[2] + w * c[2];
      C[3] = C[3] + w * c[3];
    }
    if (out) memcpy(out + 4 * i, C, sizeof C);
    if (out8) out8[i] = orc_make_rgba(C);
  }
  return 0;
}

int orc_composite_over_d(const double *imgs, const int *order, int P, long long npix,
                         double *out) {
  for (long long i = 0; i < npix; ++i) {
    double C[4] = {0, 0, 0, 0};
    for (int k = 0; k < P; ++k) {
      const double *c = imgs + ((size_t)order[k] * (size_t)npix + (size_t)i) * 4;
      const double w = 1. - C[3];
      for (int ch = 0; ch < 4; ++ch) C[ch] += w * c[ch];
    }
    memcpy(out + 4 * i, C, sizeof C);
  }
  return 0;
}

/* RGBA8 inputs (the reference composites the ranks' 8-bit images):
 * decode u/255, 'over' in float as above, re-pack. */
int orc_composite_over_u8(const uint32_t *imgs, const int *order, int P, long long npix,
                          uint32_t *out8) {
  for (long long i = 0; i < npix; ++i) {
    float C[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < P; ++k) {
      uint32_t u = imgs[(size_t)order[k] * (size_t)npix + (size_t)i];
      float c[4];
      for (int ch = 0; ch < 4; ++ch) c[ch] = (float)((u >> (8 * ch)) & 255u) / 255.f;
      const float w = 1.f - C[3];
      for (int ch = 0; ch < 4; ++ch) C[ch] = C[ch] + w * c[ch];
    }
    out8[i] = orc_make_rgba(C);
  }
  return 0;
}

/* z-buffer composite: smallest depth wins, ties to the lower rank id. */
int orc_composite_zmin(const uint32_t *imgs, const float *depths, int P, long long npix,
                       uint32_t *out8, float *outz) {
  for (long long i = 0; i < npix; ++i) {
    int best = 0;
    float bz = depths[i];
    for (int r = 1; r < P; ++r) {
      float zr = depths[(size_t)r * (size_t)npix + (size_t)i];
      if (zr < bz) { bz = zr; best = r; }
    }
    out8[i] = imgs[(size_t)best * (size_t)npix + (size_t)i];
    if (outz) outz[i] = bz;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* differentiable form: forward (float) and adjoint (float64)          */
/* ------------------------------------------------------------------ */
/* The HIP differentiable forward marches the closed-form lattice
 * t_k = dt*(j + (k0+k)) with opacity clamped to DIFF_AMAX; restated here in
 * float.  The adjoint is restated in float64 data arithmetic on the same
 * float control (same samples, same discrete decisions of the data path:
 * at a kink -- N.d = 0, a TF texel centre, the opacity clamp -- the float64
 * functions take the side the float evaluation took, so a float32 adjoint of
 * the same samples is compared with the derivative on its own branch): it
 * stores the samples of a ray and computes
 * transmittance forward as a product, then back-propagates with suffix
 * sums -- a different method from the HIP kernel's backward inversion. */
#define DIFF_AMAX 0.9999f

/* One sample of the float control: its brick, t, and the data path's
 * discrete decisions as the float march took them -- the sign of N.d (the
 * side of the |N.d| kink), the TF segment i0 and whether the opacity clamp
 * applied -- which the float64 functions below follow (as they follow the
 * float sample lattice), so that they and their derivatives sit on the same
 * branches as any float32 evaluation of the same samples. */
typedef struct {
  int brick;
  float t;
  int i0;
  signed char sgn;
  signed char clamped;
} diff_sample;

/* TF segment index of value v as tf_f takes it. */
static inline int tf_segment_f(const orc_params *p, float v) {
  const int n = p->tf_n;
  const float xt = v * (float)n - 0.5f;
  const float fi = floorf(xt);
  float fc = fi < -2.f ? -2.f : (fi > (float)n ? (float)n : fi);
  if (!(fi == fi)) fc = -2.f;
  return (int)fc;
}

/* TF lookup in double on a given segment i (w may leave [0,1] by rounding:
 * the segment's line is extended). */
static inline void tf_d_seg(const orc_params *p, double v, int i, double out[4]) {
  const int n = p->tf_n;
  const double w = (v * (double)n - 0.5) - (double)i;
  for (int c = 0; c < 4; ++c) {
    const double e0 = (i >= 0 && i < n) ? p->tf[4 * i + c] : 0.;
    const double e1 = (i + 1 >= 0 && i + 1 < n) ? p->tf[4 * (i + 1) + c] : 0.;
    out[c] = lerpd_(e0, e1, w);
  }
}

/* float control of one pixel: appends its samples (brick, t) to out (cap),
 * returns the count; fills the float colour. */
static int diff_march_pixel(const orc_params *p, const bview *bv, int nb, const float o[3],
                            const float d[3], uint32_t pixel, diff_sample *out, int cap,
                            float c[4]) {
  uint32_t rng = rng_seed(pixel, 0u);
  float tmin = 0.f;
  int n = 0;
  c[0] = c[1] = c[2] = c[3] = 0.f;
  for (;;) {
    int hit = -1;
    float best = 0.f, bt0 = 0.f, bt1 = 0.f;
    for (int i = 0; i < nb; ++i) {
      float t0, t1, rep;
      if (!slab(o, d, tmin, 1e10f, bv[i].b->space_lo, bv[i].b->space_hi, &t0, &t1)) continue;
      if (t0 > tmin) rep = t0;
      else if (t1 > tmin) rep = t1;
      else continue;
      if (hit < 0 || rep < best) { hit = i; best = rep; bt0 = t0; bt1 = t1; }
    }
    if (hit < 0) break;
    const float j = p->jitter_mode == 1 ? p->jitter_value : rng_next(&rng);
    const float dt = p->dt;
    const float segs = floorf((bt0 - dt * j) / dt);
    const float tf = dt * (j + segs);
    const float k0 = tf < bt0 ? segs + 1.f : segs;
    const orc_brick *b = bv[hit].b;
    const float vl[3] = {(float)b->vox_lo[0], (float)b->vox_lo[1], (float)b->vox_lo[2]};
    for (int k = 0;; ++k) {
      const float t = dt * (j + (k0 + (float)k));
      if (!(t < bt1 && c[3] < .99f)) break;
      diff_sample *rec = n < cap ? &out[n] : NULL;
      if (rec) { rec->brick = hit; rec->t = t; }
      ++n;
      float x = (o[0] + t * d[0]) - vl[0];
      float y = (o[1] + t * d[1]) - vl[1];
      float z = (o[2] + t * d[2]) - vl[2];
      float gx = tri_f(&bv[hit], x + .5f, y, z) - tri_f(&bv[hit], x - .5f, y, z);
      float gy = tri_f(&bv[hit], x, y + .5f, z) - tri_f(&bv[hit], x, y - .5f, z);
      float gz = tri_f(&bv[hit], x, y, z + .5f) - tri_f(&bv[hit], x, y, z - .5f);
      float ex = gx + 1e-4f, ey = gy + 1e-4f, ez = gz + 1e-4f;
      float len = sqrtf(ex * ex + ey * ey + ez * ez);
      float nx = gx / len, ny = gy / len, nz = gz / len;
      float s[4];
      const float val = tri_f(&bv[hit], x, y, z);
      tf_f(p, val, s);
      const int clamped = !(s[3] < DIFF_AMAX);
      s[3] = fminf(s[3], DIFF_AMAX);
      const float ndf = nx * d[0] + ny * d[1] + nz * d[2];
      float shade = .1f + .9f * fabsf(ndf);
      if (rec) {
        rec->i0 = tf_segment_f(p, val);
        rec->sgn = (signed char)(ndf > 0.f ? 1 : (ndf < 0.f ? -1 : 0));
        rec->clamped = (signed char)clamped;
      }
      float w = (1.f - c[3]) * s[3];
      c[0] = c[0] + (w * s[0]) * shade;
      c[1] = c[1] + (w * s[1]) * shade;
      c[2] = c[2] + (w * s[2]) * shade;
      c[3] = c[3] + w;
    }
    tmin = bt1 + 1e-20f;
    if (!(c[3] < .99f)) break;
  }
  return n;
}

static void ray_of(const orc_params *p, int px, int py, float d[3]) {
  const float su = (float)px + .5f, sv = (float)py + .5f;
  for (int a = 0; a < 3; ++a) d[a] = (p->dir00[a] + su * p->du[a]) + sv * p->dv[a];
  float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int a = 0; a < 3; ++a) d[a] = d[a] / len;
  for (int a = 0; a < 3; ++a)
    if (fabsf(d[a]) < 1e-5f) d[a] = 1e-5f;
}

static int init_views(const orc_brick *bricks, int nb, bview *bv) {
  if (nb > 256) return -2;
  for (int i = 0; i < nb; ++i) {
    bv[i].b = &bricks[i];
    for (int a = 0; a < 3; ++a) bv[i].n[a] = bricks[i].vox_hi[a] - bricks[i].vox_lo[a];
    bv[i].touched = NULL;
  }
  return 0;
}

int orc_diff_render(const orc_params *p, const orc_brick *bricks, int nb, float *rgba,
                    uint32_t *nsamples, int nthreads) {
  bview bv[256];
  if (init_views(bricks, nb, bv)) return -2;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int py = 0; py < p->H; ++py)
    for (int px = 0; px < p->W; ++px) {
      float d[3], c[4];
      ray_of(p, px, py, d);
      const size_t pix = (size_t)py * p->W + px;
      int n = diff_march_pixel(p, bv, nb, p->org, d, (uint32_t)pix, NULL, 0, c);
      memcpy(rgba + 4 * pix, c, sizeof c);
      nsamples[pix] = (uint32_t)n;
    }
  return 0;
}

static void scatter_d(double *grad, const bview *v, double x, double y, double z, double g) {
  if (g == 0.0 || !grad) return;
  double fx = floor(x), fy = floor(y), fz = floor(z);
  double wx = x - fx, wy = y - fy, wz = z - fz;
  int ix = floor_to_int((float)fx), iy = floor_to_int((float)fy), iz = floor_to_int((float)fz);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int xx = clampi(ix + i, 0, v->n[0] - 1), yy = clampi(iy + j, 0, v->n[1] - 1),
            zz = clampi(iz + k, 0, v->n[2] - 1);
        double w = (i ? wx : 1 - wx) * (j ? wy : 1 - wy) * (k ? wz : 1 - wz);
        double *dst = &grad[((size_t)zz * v->n[1] + yy) * v->n[0] + xx];
#ifdef _OPENMP
#pragma omp atomic
#endif
        *dst += g * w;
      }
}

/* dl_drgba: H*W*4 floats; dtf: tf_n*4 doubles (accumulated); dvox[i]:
 * brick i's voxel count doubles (accumulated, may be NULL). */
int orc_diff_adjoint(const orc_params *p, const orc_brick *bricks, int nb, const float *dl_drgba,
                     double *dtf, double **dvox, int nthreads) {
  bview bv[256];
  if (init_views(bricks, nb, bv)) return -2;
  const int n_tf = p->tf_n;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    int cap = 1 << 16;
    diff_sample *smp = (diff_sample *)malloc(sizeof(diff_sample) * (size_t)cap);
    double *A = (double *)malloc(sizeof(double) * (size_t)cap * 12);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int py = 0; py < p->H; ++py)
      for (int px = 0; px < p->W; ++px) {
        float d[3], c[4];
        ray_of(p, px, py, d);
        const size_t pix = (size_t)py * p->W + px;
        int n = diff_march_pixel(p, bv, nb, p->org, d, (uint32_t)pix, smp, cap, c);
        if (n == 0) continue;
        if (n > cap) n = cap; /* not reached for the sizes tested */
        const double G[4] = {dl_drgba[4 * pix], dl_drgba[4 * pix + 1], dl_drgba[4 * pix + 2],
                             dl_drgba[4 * pix + 3]};
        /* forward in double: per sample a, c(3), shade, T, plus data for backprop */
        double T = 1.0;
        for (int k = 0; k < n; ++k) {
          const bview *v = &bv[smp[k].brick];
          const double t = smp[k].t;
          double X = ((double)p->org[0] + t * d[0]) - v->b->vox_lo[0];
          double Y = ((double)p->org[1] + t * d[1]) - v->b->vox_lo[1];
          double Z = ((double)p->org[2] + t * d[2]) - v->b->vox_lo[2];
          double gx = tri_d(v, X + .5, Y, Z) - tri_d(v, X - .5, Y, Z);
          double gy = tri_d(v, X, Y + .5, Z) - tri_d(v, X, Y - .5, Z);
          double gz = tri_d(v, X, Y, Z + .5) - tri_d(v, X, Y, Z - .5);
          double L = sqrt((gx + 1e-4) * (gx + 1e-4) + (gy + 1e-4) * (gy + 1e-4) +
                          (gz + 1e-4) * (gz + 1e-4));
          double val = tri_d(v, X, Y, Z);
          double s[4];
          tf_d_seg(p, val, smp[k].i0, s);
          double araw = s[3];
          double a = smp[k].clamped ? DIFF_AMAX : araw;
          double nd = (gx / L) * d[0] + (gy / L) * d[1] + (gz / L) * d[2];
          double sh = .1 + .9 * (smp[k].sgn * nd);
          double *r = A + 12 * (size_t)k;
          r[0] = a; r[1] = s[0]; r[2] = s[1]; r[3] = s[2]; r[4] = sh; r[5] = T;
          r[6] = val; r[7] = gx; r[8] = gy; r[9] = gz; r[10] = L; r[11] = araw;
          T *= (1.0 - a);
        }
        const double TN = T;
        double S[3] = {0, 0, 0};
        for (int k = n - 1; k >= 0; --k) {
          const bview *v = &bv[smp[k].brick];
          const double *r = A + 12 * (size_t)k;
          const double a = r[0], sh = r[4], Tk = r[5], val = r[6], L = r[10];
          const double cc[3] = {r[1], r[2], r[3]};
          const double g3[3] = {r[7], r[8], r[9]};
          const double Ta = Tk * a;
          double Gcs = 0, GS = 0, Gc = 0;
          for (int q = 0; q < 3; ++q) { Gcs += G[q] * cc[q] * sh; GS += G[q] * S[q]; Gc += G[q] * cc[q]; }
          double dc[3];
          for (int q = 0; q < 3; ++q) dc[q] = G[q] * Ta * sh;
          const double dshade = Ta * Gc;
          double dalpha = Tk * Gcs - (GS - G[3] * TN) / (1.0 - a);
          for (int q = 0; q < 3; ++q) S[q] += Ta * cc[q] * sh;
          if (smp[k].clamped) dalpha = 0.0;
          /* TF entries, on the float march's segment */
          const int i0 = smp[k].i0;
          const double w = (val * n_tf - 0.5) - (double)i0;
          const double g4[4] = {dc[0], dc[1], dc[2], dalpha};
          for (int ch = 0; ch < 4; ++ch) {
            if (i0 >= 0 && i0 < n_tf) {
#ifdef _OPENMP
#pragma omp atomic
#endif
              dtf[4 * i0 + ch] += (1 - w) * g4[ch];
            }
            if (i0 + 1 >= 0 && i0 + 1 < n_tf) {
#ifdef _OPENMP
#pragma omp atomic
#endif
              dtf[4 * (i0 + 1) + ch] += w * g4[ch];
            }
          }
          double *grad = dvox ? dvox[smp[k].brick] : NULL;
          if (!grad) continue;
          double e0[4], e1[4];
          for (int ch = 0; ch < 4; ++ch) {
            e0[ch] = (i0 >= 0 && i0 < n_tf) ? p->tf[4 * i0 + ch] : 0.0;
            e1[ch] = (i0 + 1 >= 0 && i0 + 1 < n_tf) ? p->tf[4 * (i0 + 1) + ch] : 0.0;
          }
          double dv = 0;
          for (int ch = 0; ch < 4; ++ch) dv += g4[ch] * (e1[ch] - e0[ch]);
          dv *= n_tf;
          const double sg = .9 * smp[k].sgn;
          double dN[3], dNg = 0;
          for (int q = 0; q < 3; ++q) { dN[q] = dshade * sg * d[q]; dNg += dN[q] * g3[q]; }
          double dg[3];
          for (int q = 0; q < 3; ++q) dg[q] = dN[q] / L - dNg * (g3[q] + 1e-4) / (L * L * L);
          const double t = smp[k].t;
          double X = ((double)p->org[0] + t * d[0]) - v->b->vox_lo[0];
          double Y = ((double)p->org[1] + t * d[1]) - v->b->vox_lo[1];
          double Z = ((double)p->org[2] + t * d[2]) - v->b->vox_lo[2];
          scatter_d(grad, v, X, Y, Z, dv);
          scatter_d(grad, v, X + .5, Y, Z, dg[0]);
          scatter_d(grad, v, X - .5, Y, Z, -dg[0]);
          scatter_d(grad, v, X, Y + .5, Z, dg[1]);
          scatter_d(grad, v, X, Y - .5, Z, -dg[1]);
          scatter_d(grad, v, X, Y, Z + .5, dg[2]);
          scatter_d(grad, v, X, Y, Z - .5, -dg[2]);
        }
      }
    free(smp);
    free(A);
  }
  return 0;
}

/* float64 colour of the differentiable forward on the float control -- the
 * float march's samples and its discrete decisions (sign of N.d, TF
 * segment, opacity clamp) -- the function whose derivatives orc_diff_adjoint
 * computes; out: H*W*4. */
int orc_diff_render64(const orc_params *p, const orc_brick *bricks, int nb, double *out,
                      int nthreads) {
  bview bv[256];
  if (init_views(bricks, nb, bv)) return -2;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    int cap = 1 << 16;
    diff_sample *smp = (diff_sample *)malloc(sizeof(diff_sample) * (size_t)cap);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int py = 0; py < p->H; ++py)
      for (int px = 0; px < p->W; ++px) {
        float d[3], c[4];
        ray_of(p, px, py, d);
        const size_t pix = (size_t)py * p->W + px;
        int n = diff_march_pixel(p, bv, nb, p->org, d, (uint32_t)pix, smp, cap, c);
        if (n > cap) n = cap;
        double T = 1.0, C[3] = {0, 0, 0};
        for (int k = 0; k < n; ++k) {
          const bview *v = &bv[smp[k].brick];
          const double t = smp[k].t;
          double X = ((double)p->org[0] + t * d[0]) - v->b->vox_lo[0];
          double Y = ((double)p->org[1] + t * d[1]) - v->b->vox_lo[1];
          double Z = ((double)p->org[2] + t * d[2]) - v->b->vox_lo[2];
          double gx = tri_d(v, X + .5, Y, Z) - tri_d(v, X - .5, Y, Z);
          double gy = tri_d(v, X, Y + .5, Z) - tri_d(v, X, Y - .5, Z);
          double gz = tri_d(v, X, Y, Z + .5) - tri_d(v, X, Y, Z - .5);
          double L = sqrt((gx + 1e-4) * (gx + 1e-4) + (gy + 1e-4) * (gy + 1e-4) +
                          (gz + 1e-4) * (gz + 1e-4));
          double s[4];
          tf_d_seg(p, tri_d(v, X, Y, Z), smp[k].i0, s);
          double a = smp[k].clamped ? DIFF_AMAX : s[3];
          double sh = .1 + .9 * (smp[k].sgn * ((gx / L) * d[0] + (gy / L) * d[1] + (gz / L) * d[2]));
          for (int q = 0; q < 3; ++q) C[q] += T * a * s[q] * sh;
          T *= 1.0 - a;
        }
        out[4 * pix] = C[0];
        out[4 * pix + 1] = C[1];
        out[4 * pix + 2] = C[2];
        out[4 * pix + 3] = 1.0 - T;
      }
    free(smp);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* surface path: closest hit + ModelCH shading                         */
/* deviceCode.cpp:59-67 (randomColor), :176-188 (ModelCH),             */
/* :340-416 (renderFrame with traceRay(world)), :94-107 (generateRay)  */
/* ------------------------------------------------------------------ */
/* Ray/triangle test: Woop, Benthin, Wald, "Watertight Ray/Triangle
 * Intersection", JCGT 2(1) 2013 -- shear to ray space, edge functions
 * recomputed in double when one is exactly zero.  Closest hit: smallest
 * t > 0 below tmax = 1e10, ties to the smaller primitive id. */
typedef struct {
  float o[3], d[3];
  int kx, ky, kz;
  float sx, sy, sz;
} tri_ray;

static void tri_ray_setup(const float o[3], const float d[3], tri_ray *r) {
  for (int k = 0; k < 3; ++k) { r->o[k] = o[k]; r->d[k] = d[k]; }
  int kz = 0;
  if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
  if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (d[kz] < 0.f) { int t = kx; kx = ky; ky = t; }
  r->kx = kx; r->ky = ky; r->kz = kz;
  r->sx = d[kx] / d[kz];
  r->sy = d[ky] / d[kz];
  r->sz = 1.f / d[kz];
}

/* t of the hit, -1 on a miss; v: three vertices of 3 floats */
static float tri_hit(const tri_ray *r, const float *v0, const float *v1, const float *v2) {
  float A[3], B[3], Cv[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = v0[k] - r->o[k];
    B[k] = v1[k] - r->o[k];
    Cv[k] = v2[k] - r->o[k];
  }
  const float ax = A[r->kx] - r->sx * A[r->kz], ay = A[r->ky] - r->sy * A[r->kz];
  const float bx = B[r->kx] - r->sx * B[r->kz], by = B[r->ky] - r->sy * B[r->kz];
  const float cx = Cv[r->kx] - r->sx * Cv[r->kz], cy = Cv[r->ky] - r->sy * Cv[r->kz];
  float u = cx * by - cy * bx, v = ax * cy - ay * cx, w = bx * ay - by * ax;
  if (u == 0.f || v == 0.f || w == 0.f) {
    u = (float)((double)cx * (double)by - (double)cy * (double)bx);
    v = (float)((double)ax * (double)cy - (double)ay * (double)cx);
    w = (float)((double)bx * (double)ay - (double)by * (double)ax);
  }
  if ((u < 0.f || v < 0.f || w < 0.f) && (u > 0.f || v > 0.f || w > 0.f)) return -1.f;
  const float det = u + v + w;
  if (det == 0.f) return -1.f;
  const float az = r->sz * A[r->kz], bz = r->sz * B[r->kz], cz = r->sz * Cv[r->kz];
  const float T = u * az + v * bz + w * cz;
  return T / det;
}

static void pixel_dir(const orc_params *p, int px, int py, float d[3]) {
  const float su = (float)px + .5f, sv = (float)py + .5f;
  for (int a = 0; a < 3; ++a) d[a] = (p->dir00[a] + su * p->du[a]) + sv * p->dv[a];
  const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int a = 0; a < 3; ++a) d[a] = d[a] / len;
  for (int a = 0; a < 3; ++a)
    if (fabsf(d[a]) < 1e-5f) d[a] = 1e-5f;
}

/* randomColor(idx), deviceCode.cpp:59-67 */
void orc_random_color(uint32_t idx, float out[3]) {
  const uint32_t r = idx * 13u * 17u + 0x234235u;
  const uint32_t g = idx * 7u * 3u * 5u + 0x773477u;
  const uint32_t b = idx * 11u * 19u + 0x223766u;
  out[0] = (float)(r & 255u) / 255.f;
  out[1] = (float)(g & 255u) / 255.f;
  out[2] = (float)(b & 255u) / 255.f;
}

/* ModelCH's geometric normal and the raygen's headlight colour */
static void tri_shade(const float d[3], const float *v0, const float *v1, const float *v2,
                      uint32_t color, float c[4]) {
  const float e1x = v1[0] - v0[0], e1y = v1[1] - v0[1], e1z = v1[2] - v0[2];
  const float e2x = v2[0] - v0[0], e2y = v2[1] - v0[1], e2z = v2[2] - v0[2];
  const float nx = e1y * e2z - e1z * e2y;
  const float ny = e1z * e2x - e1x * e2z;
  const float nz = e1x * e2y - e1y * e2x;
  const float inv = 1.f / sqrtf(nx * nx + ny * ny + nz * nz);
  const float gx = nx * inv, gy = ny * inv, gz = nz * inv;
  const float lam = fmaxf(0.f, (-d[0]) * gx + (-d[1]) * gy + (-d[2]) * gz);
  float cd[3];
  orc_random_color(color, cd);
  for (int k = 0; k < 3; ++k) c[k] = .2f + cd[k] * lam;
  c[3] = 1.f;
}

/* per-pixel image value of a rank (accum_id 0): depth, float RGBA, RGBA8 */
static void tri_finish(const orc_params *p, int px, int py, const float c[4], float z,
                       size_t pix, float *rgba, uint32_t *rgba8, float *depth) {
  const float inv_spp = 1.f / (float)p->spp;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < p->spp; ++s)
    for (int k = 0; k < 4; ++k) acc[k] = acc[k] + c[k];
  if (depth) depth[pix] = z;
  if (p->heat) heat_blend(p, acc, NULL);
  if (rgba)
    for (int k = 0; k < 4; ++k) rgba[4 * pix + k] = (acc[k] * 1.f) * inv_spp;
  if (rgba8) {
    const int cross = p->crosshair && (px == p->W / 2 || py == p->H / 2);
    float q[4];
    for (int k = 0; k < 4; ++k) {
      float a = acc[k] * 1.f;
      if (cross) a = 1.f - a;
      q[k] = a * inv_spp;
    }
    rgba8[pix] = orc_make_rgba(q);
  }
}

static int island_pixel_of(const orc_params *p, int px, int py) {
  const int gx = (px + p->W - 1) / p->W;
  const int block = px / 16 + gx * (py / 16);
  return block % p->num_islands == p->island_id;
}

/* One rank's surface image by brute force over its triangles (tri9:
 * ntris x 9 floats in the rank's list order, primitive id = index;
 * color: colour id per triangle).  Rows [row0,row1); every output may be
 * NULL; hit_prim/hit_t: -1 / 0 for a miss.  accum_id 0. */
int orc_tri_render(const orc_params *p, const float *tri9, const uint32_t *color, long long ntris,
                   int row0, int row1, float *rgba, uint32_t *rgba8, float *depth,
                   int32_t *hit_prim, float *hit_t, int nthreads) {
  if (!p || p->W <= 0 || p->H <= 0 || p->spp < 1 || p->num_islands < 1 || ntris < 0) return -1;
  if (row0 < 0) row0 = 0;
  if (row1 > p->H) row1 = p->H;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int py = row0; py < row1; ++py)
    for (int px = 0; px < p->W; ++px) {
      const size_t pix = (size_t)py * p->W + px;
      float c[4] = {0.f, 0.f, 0.f, 0.f}, z = 1.f, best = 1e10f, d[3];
      long long hit = -1;
      if (island_pixel_of(p, px, py)) {
        tri_ray r;
        pixel_dir(p, px, py, d);
        tri_ray_setup(p->org, d, &r);
        for (long long i = 0; i < ntris; ++i) {
          const float *v = tri9 + 9 * i;
          const float t = tri_hit(&r, v, v + 3, v + 6);
          if (t > 0.f && t < best) { best = t; hit = i; }
        }
        if (hit >= 0) {
          const float *v = tri9 + 9 * hit;
          tri_shade(d, v, v + 3, v + 6, color[hit], c);
          z = best / p->max_depth;
        }
      }
      tri_finish(p, px, py, c, z, pix, rgba, rgba8, depth);
      if (hit_prim) hit_prim[pix] = (int32_t)hit;
      if (hit_t) hit_t[pix] = hit >= 0 ? best : 0.f;
    }
  return 0;
}

/* The same image by walking a BVH in the library's device layout
 * (nodes: 32 x 4-byte words each -- lo.x[4], hi.x[4], lo.y[4], hi.y[4],
 * lo.z[4], hi.z[4], child refs[4] (0x7fffffff: unused), 4 zero words;
 * tris: 12 floats each, vertex w = primitive id / colour id bits).  Used
 * where brute force is too slow; the traversal is ordered and culled
 * differently from the kernel's but returns the same closest hit (smallest
 * t, then smallest primitive id). */
int orc_tri_render_bvh(const orc_params *p, const void *nodes, const float *tris, int32_t root,
                       int row0, int row1, float *rgba, uint32_t *rgba8, float *depth,
                       int32_t *hit_prim, float *hit_t, int nthreads) {
  if (!p || p->W <= 0 || p->H <= 0 || p->spp < 1 || p->num_islands < 1) return -1;
  if (row0 < 0) row0 = 0;
  if (row1 > p->H) row1 = p->H;
  const float *nf = (const float *)nodes;
  const int32_t *ni = (const int32_t *)nodes;
  int overflow = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int py = row0; py < row1; ++py)
    for (int px = 0; px < p->W; ++px) {
      const size_t pix = (size_t)py * p->W + px;
      float c[4] = {0.f, 0.f, 0.f, 0.f}, z = 1.f, best = 1e10f, d[3];
      long long hit = -1;
      int32_t hit_id = -1;
      if (island_pixel_of(p, px, py) && root != INT_MIN) {
        tri_ray r;
        pixel_dir(p, px, py, d);
        tri_ray_setup(p->org, d, &r);
        /* double-precision slab tests widened by 1e-6 relative: culling
         * only has to be conservative */
        int32_t stack[160];
        int sp = 0;
        stack[sp++] = root;
        while (sp > 0) {
          const int32_t ref = stack[--sp];
          if (ref >= 0) {
            const float *n = nf + 32 * (size_t)ref;
            for (int k = 3; k >= 0; --k) {
              const int32_t kid = ni[32 * (size_t)ref + 24 + k];
              if (kid == 0x7fffffff) continue; /* unused slot */
              const float lo[3] = {n[k], n[8 + k], n[16 + k]};
              const float hi[3] = {n[4 + k], n[12 + k], n[20 + k]};
              double t0 = 0.0, t1 = (double)best * (1.0 + 1e-6);
              int ok = 1;
              for (int a = 0; a < 3 && ok; ++a) {
                const double inv = 1.0 / (double)d[a];
                double ta = ((double)lo[a] - p->org[a]) * inv, tb = ((double)hi[a] - p->org[a]) * inv;
                if (ta > tb) { const double t = ta; ta = tb; tb = t; }
                ta -= 1e-6 * fabs(ta) + 1e-12;
                tb += 1e-6 * fabs(tb) + 1e-12;
                if (ta > t0) t0 = ta;
                if (tb < t1) t1 = tb;
                ok = t0 <= t1;
              }
              if (ok) {
                if (sp < 160)
                  stack[sp++] = kid;
                else
                  overflow = 1;
              }
            }
          } else {
            const uint32_t leaf = ~(uint32_t)ref;
            const long long first = leaf >> 3, cnt = (leaf & 7u) + 1;
            for (long long i = first; i < first + cnt; ++i) {
              const float *t12 = tris + 12 * (size_t)i;
              int32_t prim;
              memcpy(&prim, t12 + 3, 4);
              const float t = tri_hit(&r, t12, t12 + 4, t12 + 8);
              if (t > 0.f && (t < best || (t == best && prim < hit_id))) {
                best = t;
                hit = i;
                hit_id = prim;
              }
            }
          }
        }
        if (hit >= 0) {
          const float *t12 = tris + 12 * (size_t)hit;
          uint32_t col;
          memcpy(&col, t12 + 7, 4);
          tri_shade(d, t12, t12 + 4, t12 + 8, col, c);
          z = best / p->max_depth;
        }
      }
      tri_finish(p, px, py, c, z, pix, rgba, rgba8, depth);
      if (hit_prim) hit_prim[pix] = hit_id;
      if (hit_t) hit_t[pix] = hit >= 0 ? best : 0.f;
    }
  return overflow ? -2 : 0;
}
