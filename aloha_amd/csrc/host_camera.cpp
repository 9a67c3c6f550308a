// host_camera.cpp -- screen-space bounds of bricks for the sparse launches.
//
// The raygen program sends pixel (x, y) along
//   d = clamp(normalize(dir_00 + (x+.5) du + (y+.5) dv))
// (generateRay, deviceCode.cpp:94-107), where the clamp replaces any
// component below 1e-5 in magnitude by +1e-5 (:103-105).  aloha_project_boxes
// bounds the pixels whose ray can reach a set of boxes, so the per-rank
// launch may write the miss result everywhere else without marching.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace aloha {

namespace {

struct V3 {
  double x, y, z;
};
V3 v3(const float *p) { return V3{p[0], p[1], p[2]}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 axpy(double s, V3 a, V3 b) { return V3{s * a.x + b.x, s * a.y + b.y, s * a.z + b.z}; }
V3 scale(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }

}  // namespace

// Bounds [x0,x1) x [y0,y1) of the pixels whose primary ray can reach any of
// the n boxes; the whole image when no finite bound follows.
//
// With M = [dir_00 du dv] and the rows r0, r1, r2 of its inverse, a point Q
// (relative to org) with c = M^-1 Q, c0 > 0, lies on the unclamped ray of
// pixel (c1/c0 - .5, c2/c0 - .5); a box's pixels form the convex hull of its
// corners' pixels.  The ray the kernel actually marches is n + e, with n the
// exact unit direction and |e| <= E: the clamp moves each component by less
// than 2e-5 (|e| < 2e-5 sqrt 3), and float rounding of the direction and of
// the slab test adds a few ulps of the coordinates over the distance.  Since
// r0 . n = 1/|dir_00 + (x+.5) du + (y+.5) dv| =: a, the pixel coordinate
// moves by at most (|r1| + |x+.5| |r0|) E / (a - |r0| E) in x (r2 for y).
// That bound, taken over a region reaching R pixels past the hull, is the
// margin; if it does not fit in R (or a - |r0| E <= 0 somewhere), the whole
// image is returned.
void project_boxes(const float cam[12], const aloha_box3f *boxes, int n, int W, int H,
                   int32_t out[4]) {
  out[0] = 0, out[1] = 0, out[2] = W, out[3] = H;
  if (n <= 0) {
    out[2] = out[3] = 0;
    return;
  }
  const V3 O = v3(cam), D = v3(cam + 3), U = v3(cam + 6), V = v3(cam + 9);
  const double det = dot(D, cross(U, V));
  if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) return;
  const V3 r0 = scale(1.0 / det, cross(U, V));
  const V3 r1 = scale(1.0 / det, cross(V, D));
  const V3 r2 = scale(1.0 / det, cross(D, U));
  const double n0 = norm(r0), n1 = norm(r1), n2 = norm(r2);
  double smin = 1e300, smax = -1e300, tmin = 1e300, tmax = -1e300;
  double c0min = 1e300, coord = std::max({std::fabs(O.x), std::fabs(O.y), std::fabs(O.z)});
  for (int k = 0; k < n; ++k) {
    const aloha_box3f &b = boxes[k];
    for (int c = 0; c < 8; ++c) {
      const double P[3] = {(c & 1) ? b.upper[0] : b.lower[0], (c & 2) ? b.upper[1] : b.lower[1],
                           (c & 4) ? b.upper[2] : b.lower[2]};
      if (!std::isfinite(P[0]) || !std::isfinite(P[1]) || !std::isfinite(P[2])) return;
      coord = std::max({coord, std::fabs(P[0]), std::fabs(P[1]), std::fabs(P[2])});
      const V3 Q{P[0] - O.x, P[1] - O.y, P[2] - O.z};
      const double c0 = dot(r0, Q), c1 = dot(r1, Q), c2 = dot(r2, Q);
      // corners on or behind the plane through org spanned by du, dv
      if (!(c0 > 1e-6 * norm(Q) * n0)) return;
      const double s = c1 / c0 - 0.5, t = c2 / c0 - 0.5;
      if (!std::isfinite(s) || !std::isfinite(t)) return;
      smin = std::min(smin, s), smax = std::max(smax, s);
      tmin = std::min(tmin, t), tmax = std::max(tmax, t);
      c0min = std::min(c0min, c0);
    }
  }
  // every point of the boxes lies at least c0min / |r0| from org
  const double E = 2e-5 * std::sqrt(3.0) + 1e-6 + 0x1p-19 * coord * n0 / c0min;
  const double R = 64.0;
  const double xs[2] = {smin - R + 0.5, smax + R + 0.5}, ys[2] = {tmin - R + 0.5, tmax + R + 0.5};
  double amin = 1e300;
  for (double sx : xs)
    for (double sy : ys) amin = std::min(amin, 1.0 / norm(axpy(sy, V, axpy(sx, U, D))));
  const double den = amin - n0 * E;
  if (!(den > 0.0)) return;
  const double sx_max = std::max(std::fabs(xs[0]), std::fabs(xs[1]));
  const double sy_max = std::max(std::fabs(ys[0]), std::fabs(ys[1]));
  const double mx = 2.0 + (n1 + sx_max * n0) * E / den;
  const double my = 2.0 + (n2 + sy_max * n0) * E / den;
  if (!(mx < R) || !(my < R)) return;
  const double x0 = std::floor(smin - mx), x1 = std::ceil(smax + mx) + 1.0;
  const double y0 = std::floor(tmin - my), y1 = std::ceil(tmax + my) + 1.0;
  out[0] = (int32_t)std::min<double>(W, std::max(0.0, x0));
  out[2] = (int32_t)std::min<double>(W, std::max(0.0, x1));
  out[1] = (int32_t)std::min<double>(H, std::max(0.0, y0));
  out[3] = (int32_t)std::min<double>(H, std::max(0.0, y1));
}

}  // namespace aloha

extern "C" int aloha_project_boxes(const float cam[12], const aloha_box3f *boxes, int32_t n,
                                   int32_t W, int32_t H, int32_t rect[4]) {
  if (!cam || !rect || (n > 0 && !boxes))
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_project_boxes: NULL argument");
  if (n < 0 || W < 1 || H < 1)
    return ALOHA_FAIL_GLOBAL(ALOHA_ERR_ARG, "aloha_project_boxes: n=%d, image %dx%d", n, W, H);
  aloha::project_boxes(cam, boxes, n, W, H, rect);
  return ALOHA_OK;
}
