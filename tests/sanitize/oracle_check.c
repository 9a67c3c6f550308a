/* oracle_check.c -- drives the CPU oracle (oracle/aloha_oracle.c, compiled
 * into this translation unit) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: the march with every option (u8 and f32
 * bricks, jitter, spp, accumulation, islands, crosshair, heat map, float64
 * colour, footprint marks), the composites, the differentiable forward and
 * its float64 adjoint, and the surface path by brute force and through the
 * BVH host_check wrote (the two must agree exactly).  Built by
 * tests/test_sanitize.py; argv[1]: the scratch directory holding rank.bvh. */
#include <stdio.h>

#include "../../oracle/aloha_oracle.c"

#define CHECK(c)                                                               \
  do {                                                                         \
    if (!(c)) {                                                                \
      fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

static void camera(float org[3], float d00[3], float du[3], float dv[3], int W, int H) {
  const float o[3] = {3.1f, 2.4f, -9.f}, c[3] = {-.35f, -.3f, 1.f};
  for (int k = 0; k < 3; ++k) org[k] = o[k], d00[k] = c[k];
  du[0] = .7f / W, du[1] = 0.f, du[2] = 0.f;
  dv[0] = 0.f, dv[1] = .6f / H, dv[2] = 0.f;
}

static int volume(void) {
  enum { W = 24, H = 20, NP = W * H };
  uint8_t v8[6 * 6 * 6];
  float v32[3 * 4 * 5];
  for (int i = 0; i < 216; ++i) v8[i] = (uint8_t)((i * 37) & 255);
  for (int i = 0; i < 60; ++i) v32[i] = (float)((i * 13) % 17) / 16.f;
  orc_brick b[2];
  memset(b, 0, sizeof b);
  for (int a = 0; a < 3; ++a) {
    b[0].vox_lo[a] = b[0].cell_lo[a] = 0;
    b[0].vox_hi[a] = 6, b[0].cell_hi[a] = 5;
    b[0].space_lo[a] = 0.f, b[0].space_hi[a] = 5.f;
  }
  const int n1[3] = {5, 4, 3};
  for (int a = 0; a < 3; ++a) {
    b[1].vox_lo[a] = b[1].cell_lo[a] = a == 0 ? 5 : 0;
    b[1].vox_hi[a] = b[1].cell_hi[a] = b[1].vox_lo[a] + n1[a];
    b[1].space_lo[a] = (float)b[1].vox_lo[a], b[1].space_hi[a] = (float)b[1].vox_hi[a];
  }
  b[0].dtype = ORC_U8, b[0].voxels = v8;
  b[1].dtype = ORC_F32, b[1].voxels = v32;
  float tf[4 * 16];
  for (int i = 0; i < 64; ++i) tf[i] = (float)((i * 7) % 11) / 10.f;
  orc_params p;
  memset(&p, 0, sizeof p);
  camera(p.org, p.dir00, p.du, p.dv, W, H);
  p.dt = .25f, p.max_depth = 20.f, p.W = W, p.H = H, p.spp = 2, p.accum_id = 1;
  p.island_id = 1, p.num_islands = 2, p.crosshair = 1, p.tf_n = 16, p.tf = tf;
  p.heat = 1, p.heat_scale = 1e-5f, p.heat_clock = 40000;
  static float rgba[2][NP * 4], depth[2][NP], accum[2][NP * 4];
  static uint32_t rgba8[2][NP], ns[2][NP];
  static double rgba64[NP * 4];
  uint8_t m0[216], m1[60];
  uint8_t *marks[2] = {m0, m1};
  memset(m0, 0, sizeof m0);
  memset(m1, 0, sizeof m1);
  uint64_t total = 0;
  for (int r = 0; r < 2; ++r) {
    p.jitter_mode = r;
    p.jitter_value = .3f;
    for (int i = 0; i < NP * 4; ++i) accum[r][i] = .01f * (float)(i % 7);
    CHECK(orc_render_rank(&p, b + r, 1, 0, H, rgba[r], rgba8[r], depth[r], accum[r], ns[r],
                          r ? NULL : rgba64, r ? marks + 1 : marks, 2) == 0);
    for (int i = 0; i < NP; ++i) total += ns[r][i];
  }
  CHECK(total > 0);
  /* both bricks on one rank, plain options, rows 3..9 only */
  p.spp = 1, p.accum_id = 0, p.num_islands = 1, p.island_id = 0, p.heat = 0, p.crosshair = 0;
  CHECK(orc_render_rank(&p, b, 2, 3, 9, rgba[0], rgba8[0], depth[0], NULL, ns[0], rgba64, NULL,
                        1) == 0);
  /* rejected arguments */
  orc_params bad = p;
  bad.W = 0;
  CHECK(orc_render_rank(&bad, b, 1, 0, 1, rgba[0], NULL, NULL, NULL, NULL, NULL, NULL, 1) != 0);
  CHECK(orc_render_rank(&p, b, 300, 0, 1, rgba[0], NULL, NULL, NULL, NULL, NULL, NULL, 1) != 0);
  /* composites */
  const int order[2] = {1, 0};
  static float cout[NP * 4];
  static double cout64[NP * 4], imgs64[2 * NP * 4];
  static uint32_t c8[NP];
  static float cz[NP];
  for (int i = 0; i < 2 * NP * 4; ++i) imgs64[i] = (double)((float *)rgba)[i];
  CHECK(orc_composite_over(&rgba[0][0], order, 2, NP, cout, c8) == 0);
  CHECK(orc_composite_over_d(imgs64, order, 2, NP, cout64) == 0);
  CHECK(orc_composite_over_u8(&rgba8[0][0], order, 2, NP, c8) == 0);
  CHECK(orc_composite_zmin(&rgba8[0][0], &depth[0][0], 2, NP, c8, cz) == 0);
  /* differentiable form on the f32 brick */
  orc_params q;
  memset(&q, 0, sizeof q);
  camera(q.org, q.dir00, q.du, q.dv, W, H);
  q.org[0] = 7.f;
  q.dt = .25f, q.max_depth = 1.f, q.W = W, q.H = H, q.spp = 1, q.num_islands = 1;
  q.jitter_mode = 1, q.jitter_value = .5f, q.tf_n = 16, q.tf = tf;
  static float drgba[NP * 4], dl[NP * 4];
  static uint32_t dns[NP];
  for (int i = 0; i < NP * 4; ++i) dl[i] = (float)((i % 5) - 2);
  CHECK(orc_diff_render(&q, b + 1, 1, drgba, dns, 2) == 0);
  double dtf[4 * 16], dvox[60];
  double *dv[1] = {dvox};
  memset(dtf, 0, sizeof dtf);
  memset(dvox, 0, sizeof dvox);
  CHECK(orc_diff_adjoint(&q, b + 1, 1, dl, dtf, dv, 2) == 0);
  CHECK(orc_diff_render64(&q, b + 1, 1, rgba64, 2) == 0);
  return 0;
}

static int surface(const char *dir) {
  char path[1024];
  snprintf(path, sizeof path, "%s/rank.bvh", dir);
  FILE *f = fopen(path, "rb");
  CHECK(f != NULL);
  int64_t nt = 0, nn = 0;
  int32_t root = 0;
  float cam[12];
  CHECK(fread(&nt, 8, 1, f) == 1 && fread(&nn, 8, 1, f) == 1 && fread(&root, 4, 1, f) == 1);
  CHECK(fread(cam, 4, 12, f) == 12 && nt > 0 && nt < (1 << 24) && nn >= 0 && nn < (1 << 24));
  float *nodes = (float *)malloc(128 * (size_t)(nn > 0 ? nn : 1));
  float *tris = (float *)malloc(48 * (size_t)nt);
  CHECK(fread(nodes, 128, (size_t)nn, f) == (size_t)nn && fread(tris, 48, (size_t)nt, f) == (size_t)nt);
  fclose(f);
  /* the same triangles by primitive id for the brute force */
  float *tri9 = (float *)malloc(36 * (size_t)nt);
  uint32_t *col = (uint32_t *)malloc(4 * (size_t)nt);
  for (int64_t i = 0; i < nt; ++i) {
    int32_t prim;
    memcpy(&prim, tris + 12 * i + 3, 4);
    CHECK(prim >= 0 && prim < nt);
    for (int v = 0; v < 3; ++v) memcpy(tri9 + 9 * prim + 3 * v, tris + 12 * i + 4 * v, 12);
    memcpy(col + prim, tris + 12 * i + 7, 4);
  }
  enum { W = 48, H = 40, NP = W * H };
  orc_params p;
  memset(&p, 0, sizeof p);
  memcpy(p.org, cam, 12), memcpy(p.dir00, cam + 3, 12), memcpy(p.du, cam + 6, 12);
  memcpy(p.dv, cam + 9, 12);
  p.max_depth = 4.f, p.W = W, p.H = H, p.spp = 1, p.num_islands = 1;
  static float r1[NP * 4], r2[NP * 4], z1[NP], z2[NP], t1[NP], t2[NP];
  static uint32_t a1[NP], a2[NP];
  static int32_t h1[NP], h2[NP];
  CHECK(orc_tri_render(&p, tri9, col, nt, 0, H, r1, a1, z1, h1, t1, 2) == 0);
  CHECK(orc_tri_render_bvh(&p, nodes, tris, root, 0, H, r2, a2, z2, h2, t2, 2) == 0);
  int hits = 0;
  for (int i = 0; i < NP; ++i) {
    CHECK(a1[i] == a2[i] && z1[i] == z2[i] && h1[i] == h2[i]);
    hits += h1[i] >= 0;
  }
  CHECK(hits > 0);
  free(nodes), free(tris), free(tri9), free(col);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  if (volume()) return 1;
  if (surface(argv[1])) return 1;
  printf("oracle ok\n");
  return 0;
}
