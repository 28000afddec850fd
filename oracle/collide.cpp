// ORACLE — test infrastructure only (see soft.cpp).
//
// A line-for-line restatement of the reference's narrowphase for the shapes of
// spheres and boxes (src/collide.js:204-242 collide/pair, 257-349 helpers,
// 353-403 sphere pairs, 713-1229 box against box: separating axes, face
// clipping with snap slack, feature ids, four-point reduction, edge contact,
// closest points). Pinned by the reference's own hand-worked cases
// (test/collide.test.js:61-199) and by collide() outputs the reference computed
// on seeded random pairs (tests/golden, collide.*).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "jsmath.h"

namespace {

using oracle::js_max;
using oracle::js_min;

constexpr int SPHERE = 0, CAPSULE = 1, BOX = 2, HULL = 3;
constexpr int TRIANGLE_KIND = 4;  // src/collide.js:49 (TRIANGLE)
constexpr int MAX_POINTS = 4;
constexpr double PARALLEL = 7.450580596923828e-9;  // 2^-27, src/collide.js:686

struct Hull;

struct Shape {
  int kind;
  double center[3];
  double axes[9];
  double half[3];
  double weight[3];
  const Hull* hull = nullptr;  // a hull's tables, in its body's frame
  double points[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // a triangle's corners, in the world
};

struct Manifold {
  int count;
  double normal[3];
  double points[3 * MAX_POINTS];
  double separations[MAX_POINTS];
  uint32_t ids[MAX_POINTS];
  double loss;
};

// The module scratch of src/collide.js, one call at a time.
struct Scratch {
  double unit[3];
  double seg_s, seg_t;
  double seg_a[6], seg_b[6];
  double local[6], edge[6];
  double world[3];
  double poly_a[24], poly_b[24];
  int feat_a[8], feat_b[8], line_a[8], line_b[8];
  double kept[24], kept_sep[8];
  uint32_t kept_id[8];
  double cosines[9];
  double corner_a[24], corner_b[24];
  int four[4];
  double foot[3];
  const Shape* shape_a;
  const Shape* shape_b;
  // The list reducePoints is working on (src/collide.js:1077): kept, or a hull face's.
  const double* rk;
  const double* rs;
};

const int CORNER_P[4] = {-1, 1, 1, -1}, CORNER_Q[4] = {-1, -1, 1, 1};
const int EDGE_FROM[12] = {0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3};
const int EDGE_TO[12] = {1, 3, 5, 7, 2, 3, 6, 7, 4, 5, 6, 7};

// src/collide.js:860-877: a point's feature named with the boxes' roles swapped.
struct Swapped {
  uint8_t f[32];
  Swapped() {
    const int EDGE_PLANE[4] = {3, 0, 2, 1}, PLANE_EDGE[4] = {1, 3, 2, 0};
    for (int i = 0; i < 32; i++) f[i] = (uint8_t)i;
    for (int v = 0; v < 4; v++) {
      const int j = CORNER_P[v] > 0 ? 0 : 1, k = CORNER_Q[v] > 0 ? 2 : 3;
      f[v] = (uint8_t)(20 + 4 * j + k);
      f[20 + 4 * j + k] = (uint8_t)v;
    }
    for (int e = 0; e < 4; e++)
      for (int k = 0; k < 4; k++) f[4 + 4 * e + k] = (uint8_t)(4 + 4 * PLANE_EDGE[k] + EDGE_PLANE[e]);
  }
};
const Swapped SWAPPED;

inline double sign(double x) { return x < 0 ? -1 : 1; }
inline double clamp01(double x) { return x < 0 ? 0 : x > 1 ? 1 : x; }

// src/collide.js:262-270
double normalize(Scratch& S, double x, double y, double z) {
  const double m = js_max(js_max(std::fabs(x), std::fabs(y)), std::fabs(z));
  if (m == 0) return 0;
  const double sx = x / m, sy = y / m, sz = z / m;
  const double l = std::sqrt(sx * sx + sy * sy + sz * sz);
  S.unit[0] = sx / l; S.unit[1] = sy / l; S.unit[2] = sz / l;
  return l * m;
}

void set_normal(Manifold& out, double x, double y, double z) {
  out.normal[0] = x; out.normal[1] = y; out.normal[2] = z;
}

void add_point(Manifold& out, double x, double y, double z, double separation, uint32_t id) {
  const int k = out.count++;
  out.points[3 * k] = x; out.points[3 * k + 1] = y; out.points[3 * k + 2] = z;
  out.separations[k] = separation;
  out.ids[k] = id;
}

// src/collide.js:297-306
bool balls(Scratch& S, double ax, double ay, double az, double ra, double bx, double by, double bz, double rb,
           double fx, double fy, double fz, double margin, Manifold& out, uint32_t id) {
  const double dist = normalize(S, bx - ax, by - ay, bz - az);
  const double separation = dist - ra - rb;
  if (separation > margin) return false;
  const double nx = dist > 0 ? S.unit[0] : fx, ny = dist > 0 ? S.unit[1] : fy, nz = dist > 0 ? S.unit[2] : fz;
  set_normal(out, nx, ny, nz);
  const double t = ra + separation * 0.5;
  add_point(out, ax + nx * t, ay + ny * t, az + nz * t, separation, id);
  return true;
}

// src/collide.js:326-345 (Ericson 5.1.9)
void closest_segments(Scratch& S, const double* p, const double* q) {
  const double d1x = p[3] - p[0], d1y = p[4] - p[1], d1z = p[5] - p[2];
  const double d2x = q[3] - q[0], d2y = q[4] - q[1], d2z = q[5] - q[2];
  const double rx = p[0] - q[0], ry = p[1] - q[1], rz = p[2] - q[2];
  const double a = d1x * d1x + d1y * d1y + d1z * d1z;
  const double e = d2x * d2x + d2y * d2y + d2z * d2z;
  const double f = d2x * rx + d2y * ry + d2z * rz;
  if (a == 0 && e == 0) { S.seg_s = 0; S.seg_t = 0; return; }
  if (a == 0) { S.seg_s = 0; S.seg_t = clamp01(f / e); return; }
  const double c = d1x * rx + d1y * ry + d1z * rz;
  if (e == 0) { S.seg_t = 0; S.seg_s = clamp01(-c / a); return; }
  const double b = d1x * d2x + d1y * d2y + d1z * d2z;
  const double denom = a * e - b * b;
  double s = denom != 0 ? clamp01((b * f - c * e) / denom) : 0;
  double t = (b * s + f) / e;
  if (t < 0) { t = 0; s = clamp01(-c / a); } else if (t > 1) { t = 1; s = clamp01((b - c) / a); }
  S.seg_s = s; S.seg_t = t;
}

// src/collide.js:349-353
void sphere_sphere(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  balls(S, a.center[0], a.center[1], a.center[2], a.half[0], b.center[0], b.center[1], b.center[2], b.half[0], 0, 1, 0,
        margin, out, 0);
}

// src/collide.js:367-403
void sphere_box(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  const double* ax = b.axes;
  const double* h = b.half;
  const double r = a.half[0];
  const double dx = a.center[0] - b.center[0], dy = a.center[1] - b.center[1], dz = a.center[2] - b.center[2];
  const double lx = dx * ax[0] + dy * ax[1] + dz * ax[2];
  const double ly = dx * ax[3] + dy * ax[4] + dz * ax[5];
  const double lz = dx * ax[6] + dy * ax[7] + dz * ax[8];
  const bool inside = std::fabs(lx) <= h[0] && std::fabs(ly) <= h[1] && std::fabs(lz) <= h[2];
  if (!inside) {
    const double qx = js_max(-h[0], js_min(h[0], lx));
    const double qy = js_max(-h[1], js_min(h[1], ly));
    const double qz = js_max(-h[2], js_min(h[2], lz));
    const double wx = b.center[0] + ax[0] * qx + ax[3] * qy + ax[6] * qz;
    const double wy = b.center[1] + ax[1] * qx + ax[4] * qy + ax[7] * qz;
    const double wz = b.center[2] + ax[2] * qx + ax[5] * qy + ax[8] * qz;
    balls(S, a.center[0], a.center[1], a.center[2], r, wx, wy, wz, 0, 0, 1, 0, margin, out, 0);
    return;
  }
  int axis = 0;
  double depth = h[0] - std::fabs(lx);
  if (h[1] - std::fabs(ly) < depth) { axis = 1; depth = h[1] - std::fabs(ly); }
  if (h[2] - std::fabs(lz) < depth) { axis = 2; depth = h[2] - std::fabs(lz); }
  const double l = axis == 0 ? lx : axis == 1 ? ly : lz;
  const double s = -sign(l);
  const double nx = ax[3 * axis] * s, ny = ax[3 * axis + 1] * s, nz = ax[3 * axis + 2] * s;
  set_normal(out, nx, ny, nz);
  const double separation = -(depth + r);
  const double t = r + separation * 0.5;
  add_point(out, a.center[0] + nx * t, a.center[1] + ny * t, a.center[2] + nz * t, separation, 0);
}

// src/collide.js:271-281 (perpendicular): a unit vector at right angles to v, into unit.
void perpendicular(Scratch& S, double vx, double vy, double vz) {
  const double ax = std::fabs(vx), ay = std::fabs(vy), az = std::fabs(vz);
  if (ax <= ay && ax <= az) normalize(S, 0, vz, -vy);
  else if (ay <= az) normalize(S, -vz, 0, vx);
  else normalize(S, vy, -vx, 0);
}

// src/collide.js:308-312 (segment): a capsule's core segment.
void segment(const Shape& c, double* out) {
  const double h = c.half[1];
  const double* y = c.axes;
  out[0] = c.center[0] - y[3] * h; out[1] = c.center[1] - y[4] * h; out[2] = c.center[2] - y[5] * h;
  out[3] = c.center[0] + y[3] * h; out[4] = c.center[1] + y[4] * h; out[5] = c.center[2] + y[5] * h;
}

// src/collide.js:314-319 (nearestOnSegment)
double nearest_on_segment(const double* s, double x, double y, double z) {
  const double dx = s[3] - s[0], dy = s[4] - s[1], dz = s[5] - s[2];
  const double dd = dx * dx + dy * dy + dz * dz;
  if (dd == 0) return 0;
  return clamp01(((x - s[0]) * dx + (y - s[1]) * dy + (z - s[2]) * dz) / dd);
}

// src/collide.js:357-364 (sphereCapsule)
void sphere_capsule(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  const double* c = a.center;
  segment(b, S.seg_b);
  const double* sb = S.seg_b;
  const double t = nearest_on_segment(sb, c[0], c[1], c[2]);
  const double qx = sb[0] + (sb[3] - sb[0]) * t, qy = sb[1] + (sb[4] - sb[1]) * t, qz = sb[2] + (sb[5] - sb[2]) * t;
  perpendicular(S, b.axes[3], b.axes[4], b.axes[5]);
  balls(S, c[0], c[1], c[2], a.half[0], qx, qy, qz, b.half[0], S.unit[0], S.unit[1], S.unit[2], margin, out, 0);
}

// src/collide.js:424-482 (capsuleCapsule): the closest points; for near-parallel cores the
// overlap of b's segment with a's, clipped, for two points.
void capsule_capsule(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  double* sa = S.seg_a;
  double* sb = S.seg_b;
  segment(a, sa);
  segment(b, sb);
  closest_segments(S, sa, sb);
  const double s = S.seg_s, t = S.seg_t;
  const double pax = sa[0] + (sa[3] - sa[0]) * s, pay = sa[1] + (sa[4] - sa[1]) * s, paz = sa[2] + (sa[5] - sa[2]) * s;
  const double pbx = sb[0] + (sb[3] - sb[0]) * t, pby = sb[1] + (sb[4] - sb[1]) * t, pbz = sb[2] + (sb[5] - sb[2]) * t;
  const double ra = a.half[0], rb = b.half[0];
  const double* ya = a.axes;
  const double* yb = b.axes;
  if (normalize(S, ya[4] * yb[5] - ya[5] * yb[4], ya[5] * yb[3] - ya[3] * yb[5], ya[3] * yb[4] - ya[4] * yb[3]) == 0)
    perpendicular(S, ya[3], ya[4], ya[5]);
  const double fx = S.unit[0], fy = S.unit[1], fz = S.unit[2];
  if (!balls(S, pax, pay, paz, ra, pbx, pby, pbz, rb, fx, fy, fz, margin, out, 0)) return;
  const double closest = out.separations[0];
  const double cx = out.points[0], cy = out.points[1], cz = out.points[2];
  out.count = 0;
  const double h = a.half[1];
  if (h > 0 && b.half[1] > 0) {
    const double k0 = (sb[0] - a.center[0]) * ya[3] + (sb[1] - a.center[1]) * ya[4] + (sb[2] - a.center[2]) * ya[5];
    const double k1 = (sb[3] - sb[0]) * ya[3] + (sb[4] - sb[1]) * ya[4] + (sb[5] - sb[2]) * ya[5];
    double u0 = 0, u1 = 1;
    if (k1 != 0) {
      const double ua = (-h - k0) / k1, ub = (h - k0) / k1;
      u0 = js_max(0, js_min(ua, ub));
      u1 = js_min(1, js_max(ua, ub));
    } else if (k0 < -h || k0 > h) {
      u1 = -1;
    }
    if (u1 > u0) {
      const double* n = out.normal;
      double deepest = INFINITY;
      for (int e = 0; e < 2; e++) {
        const double u = e == 0 ? u0 : u1;
        const double bx = sb[0] + (sb[3] - sb[0]) * u, by = sb[1] + (sb[4] - sb[1]) * u, bz = sb[2] + (sb[5] - sb[2]) * u;
        const double along = js_max(-h, js_min(h, k0 + k1 * u));
        const double ax = a.center[0] + ya[3] * along, ay = a.center[1] + ya[4] * along, az = a.center[2] + ya[5] * along;
        const double gx = bx - ax, gy = by - ay, gz = bz - az;
        const double normal_gap = gx * n[0] + gy * n[1] + gz * n[2];
        const double separation = normal_gap - ra - rb;
        if (separation > margin) continue;
        if (gx * gx + gy * gy + gz * gz - normal_gap * normal_gap > tolerance * tolerance) continue;
        deepest = js_min(deepest, separation);
        const double m = ra + separation * 0.5;
        add_point(out, ax + n[0] * m, ay + n[1] * m, az + n[2] * m, separation, 1 + e);
      }
      if (out.count > 0 && closest < deepest) add_point(out, cx, cy, cz, closest, 0);
      if (out.count > 0) return;
    }
  }
  add_point(out, cx, cy, cz, closest, 0);
}

void to_world(Scratch& S, const Shape& b, double x, double y, double z);

// src/collide.js:559-575 (segmentHitsBox): whether the segment (box frame) passes through the box.
bool segment_hits_box(const double* s, const double* h) {
  double t0 = 0, t1 = 1;
  for (int i = 0; i < 3; i++) {
    const double p = s[i], d = s[3 + i] - s[i];
    if (d == 0) {
      if (p < -h[i] || p > h[i]) return false;
      continue;
    }
    double ta = (-h[i] - p) / d, tb = (h[i] - p) / d;
    if (ta > tb) { const double x = ta; ta = tb; tb = x; }
    if (ta > t0) t0 = ta;
    if (tb < t1) t1 = tb;
    if (t0 > t1) return false;
  }
  return true;
}

// src/collide.js:587-618 (clipToFace): the capsule's core (S.local, box frame) clipped to
// the face `axis` on side `facing`, one or two points.
bool clip_to_face(Scratch& S, const Shape& b, int axis, double facing, double r, double margin, Manifold& out,
                  uint32_t id_base) {
  const double* h = b.half;
  const double* local = S.local;
  const int j = (axis + 1) % 3, k = (axis + 2) % 3;
  double u0 = 0, u1 = 1;
  for (int pass = 0; pass < 2; pass++) {
    const int c = pass == 0 ? j : k;
    const double p = local[c], d = local[3 + c] - local[c];
    if (d == 0) {
      if (p < -h[c] || p > h[c]) return false;
      continue;
    }
    double ta = (-h[c] - p) / d, tb = (h[c] - p) / d;
    if (ta > tb) { const double x = ta; ta = tb; tb = x; }
    if (ta > u0) u0 = ta;
    if (tb < u1) u1 = tb;
  }
  if (u0 > u1) return false;
  const double* ax = b.axes;
  set_normal(out, -facing * ax[3 * axis], -facing * ax[3 * axis + 1], -facing * ax[3 * axis + 2]);
  const int face = 2 * axis + (facing > 0 ? 1 : 0);
  for (int e = 0; e < (u1 > u0 ? 2 : 1); e++) {
    const double u = e == 0 ? u0 : u1;
    const double px = local[0] + (local[3] - local[0]) * u, py = local[1] + (local[4] - local[1]) * u,
                 pz = local[2] + (local[5] - local[2]) * u;
    const double above = facing * (axis == 0 ? px : axis == 1 ? py : pz) - h[axis];
    const double separation = above - r;
    if (separation > margin) continue;
    const double off = facing * (separation * 0.5 - above);
    to_world(S, b, axis == 0 ? px + off : px, axis == 1 ? py + off : py, axis == 2 ? pz + off : pz);
    add_point(out, S.world[0], S.world[1], S.world[2], separation, id_base + 4 * face + 1 + e);
  }
  return out.count > 0;
}

// src/collide.js:630-676 (capsuleBoxDeep): the core passes through the box; separating axes
// over the box's faces and the core crossed with its edges, a face preferred within the tolerance.
void capsule_box_deep(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  const double* h = b.half;
  const double r = a.half[0];
  const double* local = S.local;
  double* edge = S.edge;
  const double dx = local[3] - local[0], dy = local[4] - local[1], dz = local[5] - local[2];
  double best_sep = -INFINITY, nx = 0, ny = 0, nz = 0;
  int best_axis = -1;
  for (int c = 0; c < 6; c++) {
    double x = 0, y = 0, z = 0;
    if (c < 3) {
      if (c == 0) x = 1; else if (c == 1) y = 1; else z = 1;
    } else {
      const int e = c - 3;
      const double l = normalize(S, e == 0 ? 0 : e == 1 ? -dz : dy, e == 0 ? dz : e == 1 ? 0 : -dx,
                                 e == 0 ? -dy : e == 1 ? dx : 0);
      if (!(l > PARALLEL * std::sqrt(dx * dx + dy * dy + dz * dz))) continue;
      x = S.unit[0]; y = S.unit[1]; z = S.unit[2];
    }
    const double sup = h[0] * std::fabs(x) + h[1] * std::fabs(y) + h[2] * std::fabs(z);
    const double p0 = local[0] * x + local[1] * y + local[2] * z, p1 = local[3] * x + local[4] * y + local[5] * z;
    const double lo = js_min(p0, p1), hi = js_max(p0, p1);
    const double plus = lo - sup, minus = -sup - hi;
    const double sep = js_max(plus, minus) - r;
    if (sep > best_sep + (c >= 3 && best_axis < 3 ? tolerance : 0)) {
      best_sep = sep; best_axis = c;
      const double sg = plus >= minus ? -1 : 1;
      nx = x * sg; ny = y * sg; nz = z * sg;
    }
  }
  if (best_axis < 3) {
    const double facing = -(best_axis == 0 ? nx : best_axis == 1 ? ny : nz);
    clip_to_face(S, b, best_axis, sign(facing), r, margin, out, 0);
    return;
  }
  const int i = best_axis - 3, j = (i + 1) % 3, k = (i + 2) % 3;
  const double mj = j == 0 ? -nx : j == 1 ? -ny : -nz, mk = k == 0 ? -nx : k == 1 ? -ny : -nz;
  edge[i] = -h[i]; edge[3 + i] = h[i];
  edge[j] = edge[3 + j] = sign(mj) * h[j];
  edge[k] = edge[3 + k] = sign(mk) * h[k];
  closest_segments(S, local, edge);
  const double px = local[0] + dx * S.seg_s, py = local[1] + dy * S.seg_s, pz = local[2] + dz * S.seg_s;
  const double ex = edge[0] + (edge[3] - edge[0]) * S.seg_t, ey = edge[1] + (edge[4] - edge[1]) * S.seg_t,
               ez = edge[2] + (edge[5] - edge[2]) * S.seg_t;
  const double separation = (ex - px) * nx + (ey - py) * ny + (ez - pz) * nz - r;
  const double* ax = b.axes;
  set_normal(out, ax[0] * nx + ax[3] * ny + ax[6] * nz, ax[1] * nx + ax[4] * ny + ax[7] * nz,
             ax[2] * nx + ax[5] * ny + ax[8] * nz);
  const double mm = r + separation * 0.5;
  to_world(S, b, px + nx * mm, py + ny * mm, pz + nz * mm);
  add_point(out, S.world[0], S.world[1], S.world[2], separation, 32 + i);
}

// src/collide.js:497-550 (capsuleBox): the core in the box's frame; through the box, the
// deep case; else the closest of its ends and the box's twelve edges, and on a face the
// core clipped to it.
void capsule_box(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  double* sa = S.seg_a;
  double* local = S.local;
  double* edge = S.edge;
  segment(a, sa);
  const double* bx = b.axes;
  const double* h = b.half;
  const double r = a.half[0];
  for (int e = 0; e < 2; e++) {
    const double dx = sa[3 * e] - b.center[0], dy = sa[3 * e + 1] - b.center[1], dz = sa[3 * e + 2] - b.center[2];
    local[3 * e] = dx * bx[0] + dy * bx[1] + dz * bx[2];
    local[3 * e + 1] = dx * bx[3] + dy * bx[4] + dz * bx[5];
    local[3 * e + 2] = dx * bx[6] + dy * bx[7] + dz * bx[8];
  }
  if (segment_hits_box(local, h)) {
    capsule_box_deep(S, a, b, margin, tolerance, out);
    return;
  }
  double best = INFINITY, sx = 0, sy = 0, sz = 0, qx = 0, qy = 0, qz = 0;
  for (int e = 0; e < 2; e++) {
    const double px = local[3 * e], py = local[3 * e + 1], pz = local[3 * e + 2];
    const double cx = js_max(-h[0], js_min(h[0], px)), cy = js_max(-h[1], js_min(h[1], py)),
                 cz = js_max(-h[2], js_min(h[2], pz));
    const double d = (px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz);
    if (d < best) { best = d; sx = px; sy = py; sz = pz; qx = cx; qy = cy; qz = cz; }
  }
  for (int i = 0; i < 3; i++) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    for (int c = 0; c < 4; c++) {
      edge[i] = -h[i]; edge[3 + i] = h[i];
      edge[j] = edge[3 + j] = (c & 1) ? h[j] : -h[j];
      edge[k] = edge[3 + k] = (c & 2) ? h[k] : -h[k];
      closest_segments(S, local, edge);
      const double px = local[0] + (local[3] - local[0]) * S.seg_s, py = local[1] + (local[4] - local[1]) * S.seg_s,
                   pz = local[2] + (local[5] - local[2]) * S.seg_s;
      const double ex = edge[0] + (edge[3] - edge[0]) * S.seg_t, ey = edge[1] + (edge[4] - edge[1]) * S.seg_t,
                   ez = edge[2] + (edge[5] - edge[2]) * S.seg_t;
      const double d = (px - ex) * (px - ex) + (py - ey) * (py - ey) + (pz - ez) * (pz - ez);
      if (d < best) { best = d; sx = px; sy = py; sz = pz; qx = ex; qy = ey; qz = ez; }
    }
  }
  const double dist = normalize(S, qx - sx, qy - sy, qz - sz);
  if (dist - r > margin) return;
  const double nlx = S.unit[0], nly = S.unit[1], nlz = S.unit[2];
  set_normal(out, bx[0] * nlx + bx[3] * nly + bx[6] * nlz, bx[1] * nlx + bx[4] * nly + bx[7] * nlz,
             bx[2] * nlx + bx[5] * nly + bx[8] * nlz);
  const bool on_x = std::fabs(qx) == h[0], on_y = std::fabs(qy) == h[1], on_z = std::fabs(qz) == h[2];
  if ((int)on_x + (int)on_y + (int)on_z == 1) {
    const int axis = on_x ? 0 : on_y ? 1 : 2;
    const double facing = sign(axis == 0 ? qx : axis == 1 ? qy : qz);
    if (clip_to_face(S, b, axis, facing, r, margin, out, 0)) return;
  }
  const double separation = dist - r;
  const double m = r + separation * 0.5;
  to_world(S, b, sx + nlx * m, sy + nly * m, sz + nlz * m);
  add_point(out, S.world[0], S.world[1], S.world[2], separation, 0);
}

// src/collide.js:840-844
double support(const double* X, const double* h, double nx, double ny, double nz) {
  return h[0] * std::fabs(X[0] * nx + X[1] * ny + X[2] * nz) + h[1] * std::fabs(X[3] * nx + X[4] * ny + X[5] * nz)
      + h[2] * std::fabs(X[6] * nx + X[7] * ny + X[8] * nz);
}

// src/collide.js:552-557
void to_world(Scratch& S, const Shape& b, double x, double y, double z) {
  const double* ax = b.axes;
  S.world[0] = b.center[0] + ax[0] * x + ax[3] * y + ax[6] * z;
  S.world[1] = b.center[1] + ax[1] * x + ax[4] * y + ax[7] * z;
  S.world[2] = b.center[2] + ax[2] * x + ax[5] * y + ax[8] * z;
}

// src/collide.js:780-789
void corners(const Shape& x, double* c) {
  const double* X = x.axes;
  const double* h = x.half;
  for (int k = 0; k < 8; k++) {
    const double sx = (k & 1) ? h[0] : -h[0], sy = (k & 2) ? h[1] : -h[1], sz = (k & 4) ? h[2] : -h[2];
    c[3 * k] = x.center[0] + X[0] * sx + X[3] * sy + X[6] * sz;
    c[3 * k + 1] = x.center[1] + X[1] * sx + X[4] * sy + X[7] * sz;
    c[3 * k + 2] = x.center[2] + X[2] * sx + X[5] * sy + X[8] * sz;
  }
}

// src/collide.js:800-837
void box_box_closest(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  corners(a, S.corner_a);
  corners(b, S.corner_b);
  double best = INFINITY, ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  for (int side = 0; side < 2; side++) {
    const double* from = side == 0 ? S.corner_a : S.corner_b;
    const Shape& box = side == 0 ? b : a;
    const double* X = box.axes;
    const double* h = box.half;
    for (int k = 0; k < 8; k++) {
      const double dx = from[3 * k] - box.center[0], dy = from[3 * k + 1] - box.center[1],
                   dz = from[3 * k + 2] - box.center[2];
      const double lx = js_max(-h[0], js_min(h[0], dx * X[0] + dy * X[1] + dz * X[2]));
      const double ly = js_max(-h[1], js_min(h[1], dx * X[3] + dy * X[4] + dz * X[5]));
      const double lz = js_max(-h[2], js_min(h[2], dx * X[6] + dy * X[7] + dz * X[8]));
      to_world(S, box, lx, ly, lz);
      const double ex = S.world[0] - from[3 * k], ey = S.world[1] - from[3 * k + 1], ez = S.world[2] - from[3 * k + 2];
      const double d = ex * ex + ey * ey + ez * ez;
      if (d < best) {
        best = d;
        if (side == 0) {
          ax = from[3 * k]; ay = from[3 * k + 1]; az = from[3 * k + 2]; bx = S.world[0]; by = S.world[1]; bz = S.world[2];
        } else {
          bx = from[3 * k]; by = from[3 * k + 1]; bz = from[3 * k + 2]; ax = S.world[0]; ay = S.world[1]; az = S.world[2];
        }
      }
    }
  }
  for (int i = 0; i < 12; i++) {
    for (int k = 0; k < 3; k++) { S.seg_a[k] = S.corner_a[3 * EDGE_FROM[i] + k]; S.seg_a[3 + k] = S.corner_a[3 * EDGE_TO[i] + k]; }
    for (int j = 0; j < 12; j++) {
      for (int k = 0; k < 3; k++) { S.seg_b[k] = S.corner_b[3 * EDGE_FROM[j] + k]; S.seg_b[3 + k] = S.corner_b[3 * EDGE_TO[j] + k]; }
      closest_segments(S, S.seg_a, S.seg_b);
      const double* A = S.seg_a;
      const double* B = S.seg_b;
      const double px = A[0] + (A[3] - A[0]) * S.seg_s, py = A[1] + (A[4] - A[1]) * S.seg_s, pz = A[2] + (A[5] - A[2]) * S.seg_s;
      const double qx = B[0] + (B[3] - B[0]) * S.seg_t, qy = B[1] + (B[4] - B[1]) * S.seg_t, qz = B[2] + (B[5] - B[2]) * S.seg_t;
      const double d = (qx - px) * (qx - px) + (qy - py) * (qy - py) + (qz - pz) * (qz - pz);
      if (d < best) { best = d; ax = px; ay = py; az = pz; bx = qx; by = qy; bz = qz; }
    }
  }
  const double dist = normalize(S, bx - ax, by - ay, bz - az);
  if (dist > margin || dist == 0) return;
  set_normal(out, S.unit[0], S.unit[1], S.unit[2]);
  add_point(out, (ax + bx) * 0.5, (ay + by) * 0.5, (az + bz) * 0.5, dist, 1u << 14);
}

// src/collide.js:966-999 (clipPolygon, Sutherland-Hodgman with snap slack)
int clip_polygon(const double* src, const int* sf, const int* sl, int count, double* dst, int* df, int* dl, double px,
                 double py, double pz, double offset, int plane, double slack) {
  int n = 0;
  for (int v = 0; v < count; v++) {
    const int w = (v + 1) % count;
    const double ax = src[3 * v], ay = src[3 * v + 1], az = src[3 * v + 2];
    const double bx = src[3 * w], by = src[3 * w + 1], bz = src[3 * w + 2];
    const double da = ax * px + ay * py + az * pz - offset, db = bx * px + by * py + bz * pz - offset;
    const int line = sl[v];
    if (da <= slack) {
      if (da > 0) {
        dst[3 * n] = ax - px * da; dst[3 * n + 1] = ay - py * da; dst[3 * n + 2] = az - pz * da;
      } else {
        dst[3 * n] = ax; dst[3 * n + 1] = ay; dst[3 * n + 2] = az;
      }
      df[n] = sf[v];
      dl[n] = da > 0 && db > slack ? 4 + plane : line;
      n++;
    }
    if ((da <= 0 && db > slack) || (da > slack && db <= 0)) {
      const double t = da / (da - db);
      dst[3 * n] = ax + (bx - ax) * t; dst[3 * n + 1] = ay + (by - ay) * t; dst[3 * n + 2] = az + (bz - az) * t;
      df[n] = line < 4 ? 4 + 4 * line + plane
                       : 20 + 4 * (line - 4 < plane ? line - 4 : plane) + (line - 4 > plane ? line - 4 : plane);
      dl[n] = da <= 0 ? 4 + plane : line;
      n++;
    }
  }
  return n;
}

// src/collide.js:1116-1141 (four)
void four(Scratch& S, int n, int a, double nx, double ny, double nz, double pressed, double touching) {
  const double* kept = S.rk;
  const double* sep = S.rs;
  int b = -1;
  double far = -1;
  for (int pass = 0; pass < 3 && b < 0; pass++) {
    const double within = pass == 0 ? pressed : pass == 1 ? touching : INFINITY;
    for (int k = 0; k < n; k++) {
      if (k == a || sep[k] > within) continue;
      const double x = kept[3 * k] - kept[3 * a], y = kept[3 * k + 1] - kept[3 * a + 1], z = kept[3 * k + 2] - kept[3 * a + 2];
      const double d = x * x + y * y + z * z;
      if (d > far) { far = d; b = k; }
    }
  }
  int c = -1, d = -1;
  double most = 0, least = 0;
  const double abx = kept[3 * b] - kept[3 * a], aby = kept[3 * b + 1] - kept[3 * a + 1], abz = kept[3 * b + 2] - kept[3 * a + 2];
  for (int pass = 0; pass < 3 && (c < 0 || d < 0); pass++) {
    const double within = pass == 0 ? pressed : pass == 1 ? touching : INFINITY;
    const bool left = c < 0, right = d < 0;
    for (int k = 0; k < n; k++) {
      if (k == a || k == b || sep[k] > within) continue;
      const double x = kept[3 * k] - kept[3 * a], y = kept[3 * k + 1] - kept[3 * a + 1], z = kept[3 * k + 2] - kept[3 * a + 2];
      const double area = (aby * z - abz * y) * nx + (abz * x - abx * z) * ny + (abx * y - aby * x) * nz;
      if (left && area > most) { most = area; c = k; }
      if (right && area < least) { least = area; d = k; }
    }
  }
  S.four[0] = a; S.four[1] = d; S.four[2] = b; S.four[3] = c;
}

// src/collide.js:1185-1194 (foot)
bool foot(Scratch& S, const Shape& s, double nx, double ny, double nz) {
  const double* w = s.weight;
  const double* c = s.center;
  const double along = w[0] * nx + w[1] * ny + w[2] * nz;
  const int o = S.four[0];
  if (along == 0) return false;
  const double t = ((S.rk[3 * o] - c[0]) * nx + (S.rk[3 * o + 1] - c[1]) * ny + (S.rk[3 * o + 2] - c[2]) * nz) / along;
  if (!(t > 0)) return false;
  S.foot[0] = c[0] + w[0] * t; S.foot[1] = c[1] + w[1] * t; S.foot[2] = c[2] + w[2] * t;
  return true;
}

// src/collide.js:1151-1166 (rests)
bool rests(Scratch& S, const Shape& s, double nx, double ny, double nz) {
  if (!foot(S, s, nx, ny, nz)) return false;
  const double px = S.foot[0], py = S.foot[1], pz = S.foot[2];
  int from = S.four[0];
  for (int e = 1; e <= 4; e++) {
    const int to = S.four[e & 3];
    if (to < 0) continue;
    const double ax = S.rk[3 * from] - px, ay = S.rk[3 * from + 1] - py, az = S.rk[3 * from + 2] - pz;
    const double bx = S.rk[3 * to] - px, by = S.rk[3 * to + 1] - py, bz = S.rk[3 * to + 2] - pz;
    if ((ay * bz - az * by) * nx + (az * bx - ax * bz) * ny + (ax * by - ay * bx) * nz < 0) return false;
    from = to;
  }
  return true;
}

// src/collide.js:1200-1210 (under)
int under(Scratch& S, int n, double nx, double ny, double nz, double pressed) {
  int start = -1;
  double near = INFINITY;
  for (int k = 0; k < n; k++) {
    if (S.rs[k] > pressed) continue;
    const double x = S.rk[3 * k] - S.foot[0], y = S.rk[3 * k + 1] - S.foot[1], z = S.rk[3 * k + 2] - S.foot[2];
    const double along = x * nx + y * ny + z * nz;
    const double d = x * x + y * y + z * z - along * along;
    if (d < near) { near = d; start = k; }
  }
  return start;
}

// src/collide.js:1077-1104 (reducePoints), on any list of n points.
void reduce_list(Scratch& S, const double* kept, const double* sep, const uint32_t* id, int n, double nx, double ny,
                 double nz, double tolerance, Manifold& out) {
  S.rk = kept;
  S.rs = sep;
  if (n <= MAX_POINTS) {
    for (int k = 0; k < n; k++) add_point(out, kept[3 * k], kept[3 * k + 1], kept[3 * k + 2], sep[k], id[k]);
    return;
  }
  double deepest = sep[0];
  int deep = 0;
  for (int k = 1; k < n; k++)
    if (sep[k] < deepest) { deepest = sep[k]; deep = k; }
  const double pressed = js_max(deepest, 0), touching = pressed + tolerance;
  int a = -1;
  for (int k = 0; k < n; k++)
    if (sep[k] <= deepest + tolerance && (a < 0 || id[k] < id[a])) a = k;
  four(S, n, a, nx, ny, nz, pressed, touching);
  if (!rests(S, *S.shape_a, nx, ny, nz) && !rests(S, *S.shape_b, nx, ny, nz)) {
    const bool weighed = foot(S, *S.shape_a, nx, ny, nz) || foot(S, *S.shape_b, nx, ny, nz);
    bool left = false;
    for (int k = 0; k < n && !left; k++) {
      left = (weighed ? sep[k] <= pressed : k == deep) && k != S.four[0] && k != S.four[1] && k != S.four[2]
          && k != S.four[3];
    }
    if (left) four(S, n, weighed ? under(S, n, nx, ny, nz, pressed) : deep, nx, ny, nz, pressed, touching);
  }
  const int s = S.four[0], b = S.four[2], c = S.four[3], d = S.four[1];
  add_point(out, kept[3 * s], kept[3 * s + 1], kept[3 * s + 2], sep[s], id[s]);
  add_point(out, kept[3 * b], kept[3 * b + 1], kept[3 * b + 2], sep[b], id[b]);
  if (c >= 0) add_point(out, kept[3 * c], kept[3 * c + 1], kept[3 * c + 2], sep[c], id[c]);
  if (d >= 0) add_point(out, kept[3 * d], kept[3 * d + 1], kept[3 * d + 2], sep[d], id[d]);
}

// src/collide.js:1072-1074 (reduce): the box pairs' list.
void reduce_points(Scratch& S, int n, double nx, double ny, double nz, double tolerance, Manifold& out) {
  reduce_list(S, S.kept, S.kept_sep, S.kept_id, n, nx, ny, nz, tolerance, out);
}

// src/collide.js:884-945 (faceContact)
void face_contact(Scratch& S, const Shape& ref, const Shape& inc, int axis, bool flipped, double margin,
                  double tolerance, Manifold& out) {
  const double* R = ref.axes;
  const double* I = inc.axes;
  const double* hr = ref.half;
  const double* hi = inc.half;
  const double dx = inc.center[0] - ref.center[0], dy = inc.center[1] - ref.center[1], dz = inc.center[2] - ref.center[2];
  const double s = sign(dx * R[3 * axis] + dy * R[3 * axis + 1] + dz * R[3 * axis + 2]);
  const double nx = R[3 * axis] * s, ny = R[3 * axis + 1] * s, nz = R[3 * axis + 2] * s;
  const uint32_t ref_face = 2 * axis + (s > 0 ? 1 : 0);
  int inc_axis = 0;
  double best = -1;
  for (int j = 0; j < 3; j++) {
    const double d = std::fabs(I[3 * j] * nx + I[3 * j + 1] * ny + I[3 * j + 2] * nz);
    if (d > best) { best = d; inc_axis = j; }
  }
  const double inc_sign = -sign(I[3 * inc_axis] * nx + I[3 * inc_axis + 1] * ny + I[3 * inc_axis + 2] * nz);
  const uint32_t inc_face = 2 * inc_axis + (inc_sign > 0 ? 1 : 0);
  const int p = (inc_axis + 1) % 3, q = (inc_axis + 2) % 3;
  const double cx = inc.center[0] + I[3 * inc_axis] * inc_sign * hi[inc_axis];
  const double cy = inc.center[1] + I[3 * inc_axis + 1] * inc_sign * hi[inc_axis];
  const double cz = inc.center[2] + I[3 * inc_axis + 2] * inc_sign * hi[inc_axis];
  for (int v = 0; v < 4; v++) {
    const double sp = CORNER_P[v], sq = CORNER_Q[v];
    S.poly_a[3 * v] = cx + I[3 * p] * sp * hi[p] + I[3 * q] * sq * hi[q];
    S.poly_a[3 * v + 1] = cy + I[3 * p + 1] * sp * hi[p] + I[3 * q + 1] * sq * hi[q];
    S.poly_a[3 * v + 2] = cz + I[3 * p + 2] * sp * hi[p] + I[3 * q + 2] * sq * hi[q];
    S.feat_a[v] = v;
    S.line_a[v] = v;
  }
  const int u = (axis + 1) % 3, w = (axis + 2) % 3;
  int count = 4, src = 0;
  for (int plane = 0; plane < 4 && count > 0; plane++) {
    const int c = plane < 2 ? u : w;
    const double side = plane % 2 == 0 ? 1 : -1;
    const double px = R[3 * c] * side, py = R[3 * c + 1] * side, pz = R[3 * c + 2] * side;
    const double offset = px * ref.center[0] + py * ref.center[1] + pz * ref.center[2] + hr[c];
    count = src == 0 ? clip_polygon(S.poly_a, S.feat_a, S.line_a, count, S.poly_b, S.feat_b, S.line_b, px, py, pz,
                                    offset, plane, tolerance)
                     : clip_polygon(S.poly_b, S.feat_b, S.line_b, count, S.poly_a, S.feat_a, S.line_a, px, py, pz,
                                    offset, plane, tolerance);
    src = 1 - src;
  }
  const double* poly = src == 0 ? S.poly_a : S.poly_b;
  const int* feat = src == 0 ? S.feat_a : S.feat_b;
  const double fx = ref.center[0] + nx * hr[axis], fy = ref.center[1] + ny * hr[axis], fz = ref.center[2] + nz * hr[axis];
  int n = 0;
  for (int v = 0; v < count; v++) {
    const double vx = poly[3 * v], vy = poly[3 * v + 1], vz = poly[3 * v + 2];
    const double separation = (vx - fx) * nx + (vy - fy) * ny + (vz - fz) * nz;
    if (separation > margin) continue;
    const double m = separation * 0.5;
    S.kept[3 * n] = vx - nx * m; S.kept[3 * n + 1] = vy - ny * m; S.kept[3 * n + 2] = vz - nz * m;
    S.kept_sep[n] = separation;
    S.kept_id[n] = flipped ? (inc_face << 9) | (ref_face << 6) | SWAPPED.f[feat[v]]
                           : (ref_face << 9) | (inc_face << 6) | (uint32_t)feat[v];
    n++;
  }
  if (n == 0) return;
  if (flipped) set_normal(out, -nx, -ny, -nz);
  else set_normal(out, nx, ny, nz);
  reduce_points(S, n, nx, ny, nz, tolerance, out);
}

// src/collide.js:1232-1243 (boxEdge)
void box_edge(const Shape& x, int i, double nx, double ny, double nz, double* s) {
  const double* X = x.axes;
  const double* h = x.half;
  double cx = x.center[0], cy = x.center[1], cz = x.center[2];
  for (int k = 0; k < 3; k++) {
    if (k == i) continue;
    const double t = sign(X[3 * k] * nx + X[3 * k + 1] * ny + X[3 * k + 2] * nz) * h[k];
    cx += X[3 * k] * t; cy += X[3 * k + 1] * t; cz += X[3 * k + 2] * t;
  }
  const double ex = X[3 * i] * h[i], ey = X[3 * i + 1] * h[i], ez = X[3 * i + 2] * h[i];
  s[0] = cx - ex; s[1] = cy - ey; s[2] = cz - ez;
  s[3] = cx + ex; s[4] = cy + ey; s[5] = cz + ez;
}

// src/collide.js:1213-1229 (edgeContact)
void edge_contact(Scratch& S, const Shape& a, const Shape& b, int i, int j, double nx, double ny, double nz,
                  double margin, Manifold& out) {
  box_edge(a, i, nx, ny, nz, S.seg_a);
  box_edge(b, j, -nx, -ny, -nz, S.seg_b);
  closest_segments(S, S.seg_a, S.seg_b);
  const double* A = S.seg_a;
  const double* B = S.seg_b;
  const double pax = A[0] + (A[3] - A[0]) * S.seg_s, pay = A[1] + (A[4] - A[1]) * S.seg_s, paz = A[2] + (A[5] - A[2]) * S.seg_s;
  const double pbx = B[0] + (B[3] - B[0]) * S.seg_t, pby = B[1] + (B[4] - B[1]) * S.seg_t, pbz = B[2] + (B[5] - B[2]) * S.seg_t;
  const double separation = (pbx - pax) * nx + (pby - pay) * ny + (pbz - paz) * nz;
  if (separation > margin) return;
  if (separation > 0 && (S.seg_s == 0 || S.seg_s == 1 || S.seg_t == 0 || S.seg_t == 1)) return;
  set_normal(out, nx, ny, nz);
  add_point(out, (pax + pbx) * 0.5, (pay + pby) * 0.5, (paz + pbz) * 0.5, separation, (1u << 13) | (uint32_t)(3 * i + j));
}

// src/collide.js:713-775 (boxBox)
void box_box(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  const double* A = a.axes;
  const double* B = b.axes;
  const double* ha = a.half;
  const double* hb = b.half;
  const double dx = b.center[0] - a.center[0], dy = b.center[1] - a.center[1], dz = b.center[2] - a.center[2];
  double* R = S.cosines;
  double face_a = -INFINITY, face_b = -INFINITY;
  int axis_a = 0, axis_b = 0;
  for (int i = 0; i < 3; i++) {
    const double nx = A[3 * i], ny = A[3 * i + 1], nz = A[3 * i + 2];
    const double r0 = std::fabs(B[0] * nx + B[1] * ny + B[2] * nz);
    const double r1 = std::fabs(B[3] * nx + B[4] * ny + B[5] * nz);
    const double r2 = std::fabs(B[6] * nx + B[7] * ny + B[8] * nz);
    R[3 * i] = r0; R[3 * i + 1] = r1; R[3 * i + 2] = r2;
    const double s = std::fabs(dx * nx + dy * ny + dz * nz) - ha[i] - (hb[0] * r0 + hb[1] * r1 + hb[2] * r2);
    if (s > margin) return;
    if (s > face_a) { face_a = s; axis_a = i; }
  }
  for (int j = 0; j < 3; j++) {
    const double nx = B[3 * j], ny = B[3 * j + 1], nz = B[3 * j + 2];
    const double s = std::fabs(dx * nx + dy * ny + dz * nz) - hb[j] - (ha[0] * R[j] + ha[1] * R[3 + j] + ha[2] * R[6 + j]);
    if (s > margin) return;
    if (s > face_b) { face_b = s; axis_b = j; }
  }
  double edge_sep = -INFINITY, enx = 0, eny = 0, enz = 0;
  int edge_i = -1, edge_j = -1;
  for (int i = 0; i < 3; i++) {
    const double ux = A[3 * i], uy = A[3 * i + 1], uz = A[3 * i + 2];
    for (int j = 0; j < 3; j++) {
      const double vx = B[3 * j], vy = B[3 * j + 1], vz = B[3 * j + 2];
      const double l = normalize(S, uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx);
      if (!(l > PARALLEL)) continue;
      double nx = S.unit[0], ny = S.unit[1], nz = S.unit[2];
      const double d = dx * nx + dy * ny + dz * nz;
      if (d < 0) { nx = -nx; ny = -ny; nz = -nz; }
      const double s = std::fabs(d) - support(A, ha, nx, ny, nz) - support(B, hb, nx, ny, nz);
      if (s > margin) return;
      if (s > edge_sep) { edge_sep = s; edge_i = i; edge_j = j; enx = nx; eny = ny; enz = nz; }
    }
  }
  if (edge_i >= 0 && edge_sep > js_max(face_a, face_b) + tolerance) {
    edge_contact(S, a, b, edge_i, edge_j, enx, eny, enz, margin, out);
  } else if (face_b > face_a + tolerance) {
    face_contact(S, b, a, axis_b, true, margin, tolerance, out);
  } else {
    face_contact(S, a, b, axis_a, false, margin, tolerance, out);
  }
  if (out.count == 0 && js_max(js_max(face_a, face_b), edge_sep) > 0) box_box_closest(S, a, b, margin, out);
}

// ---- hulls: src/hull.js:36-131 (the tables), src/collide.js:1262-1882 (the pairs) ----

// A hull's tables in its own frame, with what src/hull.js's constructor derives from them:
// each edge once (from the first face to meet it), its two faces, the box round the
// corners, the reach, the thinnest size and the inscribed ball.
struct Hull {
  int V = 0, F = 0, E = 0;
  std::vector<double> vertices, planes;  // 3V; 4F: unit outward normal, distance
  std::vector<int> start, loops;         // F + 1; corners counter-clockwise from outside
  std::vector<int> edges, sides;         // 2E each
  double lo[3], hi[3], reach = 0, size = 0, inner = 0;
};

// src/hull.js:45-80 (the constructor's derived tables).
void derive(Hull& h) {
  h.F = (int)h.start.size() - 1;
  h.V = (int)h.vertices.size() / 3;
  const int V = h.V;
  std::vector<int> owner((size_t)V * V, -1);
  h.edges.clear();
  h.sides.clear();
  for (int f = 0; f < h.F; f++) {
    for (int i = h.start[f]; i < h.start[f + 1]; i++) {
      const int u = h.loops[i], v = h.loops[i + 1 < h.start[f + 1] ? i + 1 : h.start[f]];
      const int back = owner[(size_t)v * V + u];
      if (back < 0) owner[(size_t)u * V + v] = f;
      else { h.edges.push_back(v); h.edges.push_back(u); h.sides.push_back(back); h.sides.push_back(f); }
    }
  }
  h.E = (int)h.edges.size() / 2;
  const double* v = h.vertices.data();
  for (int k = 0; k < 3; k++) { h.lo[k] = INFINITY; h.hi[k] = -INFINITY; }
  double reach = 0;
  for (int i = 0; i < V; i++) {
    const double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    h.lo[0] = js_min(h.lo[0], x); h.lo[1] = js_min(h.lo[1], y); h.lo[2] = js_min(h.lo[2], z);
    h.hi[0] = js_max(h.hi[0], x); h.hi[1] = js_max(h.hi[1], y); h.hi[2] = js_max(h.hi[2], z);
    reach = js_max(reach, std::sqrt(x * x + y * y + z * z));
  }
  h.reach = reach;
  const double* p = h.planes.data();
  double size = INFINITY;
  for (int f = 0; f < h.F; f++) {
    const double nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
    double low = INFINITY;
    for (int i = 0; i < V; i++) low = js_min(low, nx * v[3 * i] + ny * v[3 * i + 1] + nz * v[3 * i + 2]);
    size = js_min(size, p[4 * f + 3] - low);
  }
  h.size = size;
  double inner = INFINITY;
  for (int f = 0; f < h.F; f++) {
    const double nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
    for (int i = h.start[f]; i < h.start[f + 1]; i++) {
      const int c = 3 * h.loops[i];
      inner = js_min(inner, nx * v[c] + ny * v[c + 1] + nz * v[c + 2]);
    }
  }
  h.inner = inner;
}

// src/collide.js:1298-1299 (CUBE): src/hull.js build() of the unit cube's corners, as the
// reference's build() lays it out (corners in input order, faces by the order quickhull
// finished them).
const Hull& cube() {
  static Hull h = [] {
    Hull c;
    c.vertices = {-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5,
                  -0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5};
    c.start = {0, 4, 8, 12, 16, 20, 24};
    c.loops = {0, 1, 5, 4, 0, 4, 6, 2, 0, 2, 3, 1, 1, 3, 7, 5, 4, 5, 7, 6, 2, 6, 7, 3};
    c.planes = {0, -1, 0, 0.5, -1, 0, 0, 0.5, 0, 0, -1, 0.5, 1, 0, 0, 0.5, 0, 0, 1, 0.5, 0, 1, 0, 0.5};
    derive(c);
    return c;
  }();
  return h;
}

constexpr double LOOSE = 1.0 / 65536;     // 2^-16, src/collide.js:1270
constexpr double NEARLY = 1.0 / 67108864; // 2^-26, src/collide.js:1278
constexpr int CUT = 1 << 20;              // src/collide.js:1541

// src/collide.js:1286-1295 (reaches)
bool reaches(double a1, double a2, double c, double bound) {
  const double hi = a1 > a2 ? a1 : a2, lo = a1 > a2 ? a2 : a1;
  if (!(hi < bound)) return true;
  const double s = 1 - c * c;
  if (!(s >= NEARLY)) return true;
  const double u = lo - c * hi;
  return u >= 0 && !(std::sqrt(hi * hi + u * u / s) < bound);
}

// src/collide.js:1309-1388 (Placed): a hull, or a box as the cube scaled, placed in the world.
struct Placed {
  const Hull* hull = nullptr;
  std::vector<double> v, planes, arcs;
  double center[3], inner = 0, reach = 0;

  void set(const Shape& s) {
    const bool box = s.kind == BOX;
    const Hull& h = box ? cube() : *s.hull;
    hull = &h;
    const int V = h.V, F = h.F, E = h.E;
    v.assign(3 * V, 0);
    planes.assign(4 * F, 0);
    arcs.assign(3 * E, 0);
    const double* A = s.axes;
    const double* c = s.center;
    const double* half = s.half;
    center[0] = c[0]; center[1] = c[1]; center[2] = c[2];
    inner = box ? js_min(js_min(half[0], half[1]), half[2]) : h.inner;
    reach = box ? std::sqrt(half[0] * half[0] + half[1] * half[1] + half[2] * half[2]) : h.reach;
    for (int i = 0; i < V; i++) {
      double x = h.vertices[3 * i], y = h.vertices[3 * i + 1], z = h.vertices[3 * i + 2];
      if (box) { x *= 2 * half[0]; y *= 2 * half[1]; z *= 2 * half[2]; }
      v[3 * i] = c[0] + A[0] * x + A[3] * y + A[6] * z;
      v[3 * i + 1] = c[1] + A[1] * x + A[4] * y + A[7] * z;
      v[3 * i + 2] = c[2] + A[2] * x + A[5] * y + A[8] * z;
    }
    for (int f = 0; f < F; f++) {
      const double x = h.planes[4 * f], y = h.planes[4 * f + 1], z = h.planes[4 * f + 2];
      const double nx = A[0] * x + A[3] * y + A[6] * z, ny = A[1] * x + A[4] * y + A[7] * z, nz = A[2] * x + A[5] * y + A[8] * z;
      const double d = box ? std::fabs(x) * half[0] + std::fabs(y) * half[1] + std::fabs(z) * half[2] : h.planes[4 * f + 3];
      planes[4 * f] = nx; planes[4 * f + 1] = ny; planes[4 * f + 2] = nz;
      planes[4 * f + 3] = d + nx * c[0] + ny * c[1] + nz * c[2];
    }
    const double* p = planes.data();
    for (int e = 0; e < E; e++) {
      const int l = 4 * h.sides[2 * e], r = 4 * h.sides[2 * e + 1];
      arcs[3 * e] = p[r + 1] * p[l + 2] - p[r + 2] * p[l + 1];
      arcs[3 * e + 1] = p[r + 2] * p[l] - p[r] * p[l + 2];
      arcs[3 * e + 2] = p[r] * p[l + 1] - p[r + 1] * p[l];
    }
  }

  double lowest(double nx, double ny, double nz) const {
    double m = INFINITY;
    for (int i = 0; i < hull->V; i++) {
      const double d = v[3 * i] * nx + v[3 * i + 1] * ny + v[3 * i + 2] * nz;
      if (d < m) m = d;
    }
    return m;
  }
};

// The hulls shape records name by number (a hull's record holds its number in half[0]),
// set by oracle_hulls; and the last oracle_collide's losses.
std::vector<Hull> g_hulls;
std::vector<double> g_loss;

void resolve(Shape& s) {
  s.hull = s.kind == HULL ? &g_hulls.at((size_t)s.half[0]) : nullptr;
}

// src/collide.js:1401-1404 (mix): a point's id from the features that made it.
uint32_t mix(uint32_t kind, uint32_t i, uint32_t j, uint32_t k) {
  return ((kind + 1) * 0x9e3779b1u) ^ ((i + 1) * 0x85ebca77u) ^ ((j + 1) * 0xc2b2ae3du) ^ ((k + 1) * 0x27d4eb2fu);
}

double gjk_closest(const Shape& a, const Shape& b, double* near_a, double* near_b);

// src/collide.js:1751-1759 (nearest): apart, the cores' closest points by GJK, one point.
double nearest(Scratch& S, const Shape& a, const Shape& b, double ra, double margin, Manifold& out) {
  double na[3], nb[3];
  const double d = gjk_closest(a, b, na, nb);
  if (!(d > 0) || d - ra > margin) return d;
  const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
  const double separation = d - ra, t = ra + separation * 0.5;
  set_normal(out, nx, ny, nz);
  add_point(out, na[0] + nx * t, na[1] + ny * t, na[2] + nz * t, separation, mix(3, 0, 0, 0));
  return d;
}

// src/collide.js:1570-1581 (named): a clipped point's id from the features of a and of b.
uint32_t named(int f, bool flipped, int rf, int inc) {
  int g;
  if (f < CUT) g = flipped ? f : CUT + f;
  else if (f < 2 * CUT) {
    const int x = (f - CUT) >> 1, line = x >> 12, plane = x & 4095;
    g = 2 * CUT + (flipped ? 4096 * line + plane : 4096 * plane + line);
  } else {
    const int x = (f - 2 * CUT) >> 1, p = x >> 12, q = x & 4095, c = q == p + 1 ? q : p;
    g = flipped ? CUT + c : c;
  }
  return flipped ? mix(0, inc, rf, g) : mix(0, rf, inc, g);
}

// src/collide.js:1698-1725 (clipSides): the polygon clipped to (x, y, z).m <= offset.
int clip_sides(const double* src, const int* sf, const int* sl, int count, double* dst, int* df, int* dl, double mx,
               double my, double mz, double offset, int plane, double slack, double length) {
  int n = 0;
  for (int v = 0; v < count; v++) {
    const int w = v + 1 < count ? v + 1 : 0;
    const double ax = src[3 * v], ay = src[3 * v + 1], az = src[3 * v + 2];
    const double bx = src[3 * w], by = src[3 * w + 1], bz = src[3 * w + 2];
    const double da = ax * mx + ay * my + az * mz - offset, db = bx * mx + by * my + bz * mz - offset;
    const int line = sl[v];
    if (da <= slack) {
      if (da > 0) {
        const double t = da / (length * length);
        dst[3 * n] = ax - mx * t; dst[3 * n + 1] = ay - my * t; dst[3 * n + 2] = az - mz * t;
      } else {
        dst[3 * n] = ax; dst[3 * n + 1] = ay; dst[3 * n + 2] = az;
      }
      df[n] = sf[v];
      dl[n] = da > 0 && db > slack ? CUT + plane : line;
      n++;
    }
    if ((da <= 0 && db > slack) || (da > slack && db <= 0)) {
      const double t = da / (da - db);
      dst[3 * n] = ax + (bx - ax) * t; dst[3 * n + 1] = ay + (by - ay) * t; dst[3 * n + 2] = az + (bz - az) * t;
      df[n] = line < CUT ? CUT + 2 * (line * 4096 + plane)
                         : 2 * CUT + 2 * ((line - CUT < plane ? line - CUT : plane) * 4096 + (line - CUT > plane ? line - CUT : plane));
      dl[n] = da <= 0 ? CUT + plane : line;
      n++;
    }
  }
  return n;
}

// src/collide.js:1592-1689 (hullFace): face rf of R as the reference against I's face most
// nearly opposite it, clipped; whether it added points.
bool hull_face(Scratch& S, const Placed& R, int rf, const Placed& I, bool flipped, double margin, double tolerance,
               Manifold& out, bool sure) {
  const double* rp = R.planes.data();
  const double* ip = I.planes.data();
  const Hull& rh = *R.hull;
  const Hull& ih = *I.hull;
  const double nx = rp[4 * rf], ny = rp[4 * rf + 1], nz = rp[4 * rf + 2], d = rp[4 * rf + 3];
  const double* iv = I.v.data();
  int low = 0;
  double lowest = INFINITY;
  for (int i = 0; i < ih.V; i++) {
    const double dot = iv[3 * i] * nx + iv[3 * i + 1] * ny + iv[3 * i + 2] * nz;
    if (dot < lowest) { lowest = dot; low = i; }
  }
  int inc = 0;
  double least = INFINITY;
  for (int f = 0; f < ih.F; f++) {
    const double dot = ip[4 * f] * nx + ip[4 * f + 1] * ny + ip[4 * f + 2] * nz;
    if (!(dot < least)) continue;
    bool at = false;
    for (int j = ih.start[f]; j < ih.start[f + 1] && !at; j++) at = ih.loops[j] == low;
    if (at) { least = dot; inc = f; }
  }
  const int i0 = ih.start[inc], m = ih.start[inc + 1] - i0, r0 = rh.start[rf], k = rh.start[rf + 1] - r0;
  const int cap = m + k + 1;
  std::vector<double> clip_p(3 * cap), clip_q(3 * cap), kept_p(3 * cap), kept_s(cap);
  std::vector<int> feat_p(cap), feat_q(cap), line_p(cap), line_q(cap);
  std::vector<uint32_t> kept_i(cap);
  const double* rv = R.v.data();
  for (int j = 0; j < m; j++) {
    const int v = 3 * ih.loops[i0 + j];
    clip_p[3 * j] = iv[v]; clip_p[3 * j + 1] = iv[v + 1]; clip_p[3 * j + 2] = iv[v + 2];
    feat_p[j] = j; line_p[j] = j;
  }
  int count = m, src = 0;
  for (int j = 0; j < k && count > 0; j++) {
    const int p = 3 * rh.loops[r0 + j], q = 3 * rh.loops[r0 + (j + 1 < k ? j + 1 : 0)];
    const double ex = rv[q] - rv[p], ey = rv[q + 1] - rv[p + 1], ez = rv[q + 2] - rv[p + 2];
    const double mx = ey * nz - ez * ny, my = ez * nx - ex * nz, mz = ex * ny - ey * nx;
    const double offset = mx * rv[p] + my * rv[p + 1] + mz * rv[p + 2];
    const double length = std::sqrt(mx * mx + my * my + mz * mz);
    count = src == 0 ? clip_sides(clip_p.data(), feat_p.data(), line_p.data(), count, clip_q.data(), feat_q.data(),
                                  line_q.data(), mx, my, mz, offset, j, tolerance * length, length)
                     : clip_sides(clip_q.data(), feat_q.data(), line_q.data(), count, clip_p.data(), feat_p.data(),
                                  line_p.data(), mx, my, mz, offset, j, tolerance * length, length);
    src = 1 - src;
  }
  const double* poly = src == 0 ? clip_p.data() : clip_q.data();
  const int* feat = src == 0 ? feat_p.data() : feat_q.data();
  int n = 0;
  for (int j = 0; j < count; j++) {
    const double x = poly[3 * j], y = poly[3 * j + 1], z = poly[3 * j + 2];
    const double separation = x * nx + y * ny + z * nz - d;
    if (separation > margin) continue;
    const double h = separation * 0.5;
    kept_p[3 * n] = x - nx * h; kept_p[3 * n + 1] = y - ny * h; kept_p[3 * n + 2] = z - nz * h;
    kept_s[n] = separation;
    kept_i[n] = named(feat[j], flipped, rf, inc);
    n++;
  }
  const double s = lowest - d;
  if (s <= tolerance) {
    double deepest = INFINITY;
    for (int j = 0; j < n; j++)
      if (kept_s[j] < deepest) deepest = kept_s[j];
    if (deepest > s + tolerance) {
      if (!sure) return false;
      int j = 0;
      while (ih.loops[i0 + j] != low) j++;
      const int v = 3 * low;
      const double h = s * 0.5;
      kept_p[3 * n] = iv[v] - nx * h; kept_p[3 * n + 1] = iv[v + 1] - ny * h; kept_p[3 * n + 2] = iv[v + 2] - nz * h;
      kept_s[n] = s;
      kept_i[n] = named(j, flipped, rf, inc);
      n++;
    }
  }
  if (n == 0) return false;
  if (flipped) set_normal(out, -nx, -ny, -nz);
  else set_normal(out, nx, ny, nz);
  reduce_list(S, kept_p.data(), kept_s.data(), kept_i.data(), n, nx, ny, nz, tolerance, out);
  if (s <= tolerance) {
    double kept = INFINITY;
    for (int j = 0; j < out.count; j++)
      if (out.separations[j] < kept) kept = out.separations[j];
    if (kept - s > out.loss) out.loss = kept - s;
  }
  return true;
}

// src/collide.js:1728-1744 (hullEdges): edge i of A against edge j of B, the one point.
void hull_edges(Scratch& S, const Placed& A, int i, const Placed& B, int j, double nx, double ny, double nz,
                double margin, Manifold& out) {
  const int u = 3 * A.hull->edges[2 * i], w = 3 * A.hull->edges[2 * i + 1];
  const int p = 3 * B.hull->edges[2 * j], q = 3 * B.hull->edges[2 * j + 1];
  double* sa = S.seg_a;
  double* sb = S.seg_b;
  for (int k = 0; k < 3; k++) {
    sa[k] = A.v[u + k]; sa[3 + k] = A.v[w + k];
    sb[k] = B.v[p + k]; sb[3 + k] = B.v[q + k];
  }
  closest_segments(S, sa, sb);
  const double ss = S.seg_s, tt = S.seg_t;
  const double pax = sa[0] + (sa[3] - sa[0]) * ss, pay = sa[1] + (sa[4] - sa[1]) * ss, paz = sa[2] + (sa[5] - sa[2]) * ss;
  const double pbx = sb[0] + (sb[3] - sb[0]) * tt, pby = sb[1] + (sb[4] - sb[1]) * tt, pbz = sb[2] + (sb[5] - sb[2]) * tt;
  const double separation = (pbx - pax) * nx + (pby - pay) * ny + (pbz - paz) * nz;
  if (separation > margin) return;
  if (separation > 0 && (ss == 0 || ss == 1 || tt == 0 || tt == 1)) return;
  set_normal(out, nx, ny, nz);
  add_point(out, (pax + pbx) * 0.5, (pay + pby) * 0.5, (paz + pbz) * 0.5, separation, mix(2, i, j, 0));
}

// src/collide.js:1406-1536 (hullHull): a hull, or a box as the cube, against a hull.
void hull_hull(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  Placed A, B;
  A.set(a);
  B.set(b);
  const Hull& ha = *A.hull;
  const Hull& hb = *B.hull;
  const double* pa = A.planes.data();
  const double* pb = B.planes.data();
  const double* ca = A.center;
  const double* cb = B.center;
  const int FA = ha.F, FB = hb.F;
  const double loose = LOOSE * (std::fabs(ca[0]) + std::fabs(ca[1]) + std::fabs(ca[2]) + std::fabs(cb[0]) + std::fabs(cb[1])
                                + std::fabs(cb[2]) + A.reach + B.reach);
  const double ra = A.inner > loose ? A.inner - loose : -INFINITY, rb = B.inner > loose ? B.inner - loose : -INFINITY;
  std::vector<double> above_a(FA), above_b(FB);
  int top = 0;
  for (int f = 0; f < FA; f++) {
    above_a[f] = pa[4 * f] * cb[0] + pa[4 * f + 1] * cb[1] + pa[4 * f + 2] * cb[2] - pa[4 * f + 3];
    if (above_a[f] > above_a[top]) top = f;
  }
  const double first_a = B.lowest(pa[4 * top], pa[4 * top + 1], pa[4 * top + 2]) - pa[4 * top + 3];
  if (first_a > margin) return;
  double face_a = -INFINITY;
  int fa = 0;
  for (int f = 0; f < FA; f++) {
    double sep = first_a;
    if (f != top) {
      if (above_a[f] - rb < js_max(face_a, first_a)) continue;
      sep = B.lowest(pa[4 * f], pa[4 * f + 1], pa[4 * f + 2]) - pa[4 * f + 3];
      if (sep > margin) return;
    }
    if (sep > face_a) { face_a = sep; fa = f; }
  }
  top = 0;
  for (int f = 0; f < FB; f++) {
    above_b[f] = pb[4 * f] * ca[0] + pb[4 * f + 1] * ca[1] + pb[4 * f + 2] * ca[2] - pb[4 * f + 3];
    if (above_b[f] > above_b[top]) top = f;
  }
  const double first_b = A.lowest(pb[4 * top], pb[4 * top + 1], pb[4 * top + 2]) - pb[4 * top + 3];
  if (first_b > margin) return;
  double face_b = -INFINITY;
  int fb = 0;
  for (int f = 0; f < FB; f++) {
    double sep = first_b;
    if (f != top) {
      if (above_b[f] - ra < js_max(js_max(face_b, first_b), face_a)) continue;
      sep = A.lowest(pb[4 * f], pb[4 * f + 1], pb[4 * f + 2]) - pb[4 * f + 3];
      if (sep > margin) return;
    }
    if (sep > face_b) { face_b = sep; fb = f; }
  }
  const int* ea = ha.edges.data();
  const int* sa = ha.sides.data();
  const int* eb = hb.edges.data();
  const int* sb = hb.sides.data();
  const double* va = A.v.data();
  const double* vb = B.v.data();
  const double* arc_a = A.arcs.data();
  const double* arc_b = B.arcs.data();
  const int EA = ha.E, EB = hb.E;
  std::vector<double> side(FB);
  const double best = js_max(face_a, face_b);
  std::vector<uint8_t> reach_b(EB);
  int reachable = 0;
  for (int j = 0; j < EB; j++) {
    const int l = 4 * sb[2 * j], r = 4 * sb[2 * j + 1], p = 3 * eb[2 * j];
    const double dx = ca[0] - vb[p], dy = ca[1] - vb[p + 1], dz = ca[2] - vb[p + 2];
    const bool can = reaches(pb[l] * dx + pb[l + 1] * dy + pb[l + 2] * dz, pb[r] * dx + pb[r + 1] * dy + pb[r + 2] * dz,
                             pb[l] * pb[r] + pb[l + 1] * pb[r + 1] + pb[l + 2] * pb[r + 2], best + ra);
    reach_b[j] = can ? 1 : 0;
    if (can) reachable++;
  }
  double edge_sep = -INFINITY, enx = 0, eny = 0, enz = 0;
  int ei = -1, ej = -1;
  for (int i = 0; i < EA && reachable > 0; i++) {
    const int l = 4 * sa[2 * i], r = 4 * sa[2 * i + 1];
    const int u = 3 * ea[2 * i], w = 3 * ea[2 * i + 1];
    const double dx = cb[0] - va[u], dy = cb[1] - va[u + 1], dz = cb[2] - va[u + 2];
    if (!reaches(pa[l] * dx + pa[l + 1] * dy + pa[l + 2] * dz, pa[r] * dx + pa[r + 1] * dy + pa[r + 2] * dz,
                 pa[l] * pa[r] + pa[l + 1] * pa[r + 1] + pa[l + 2] * pa[r + 2], best + rb))
      continue;
    const double ax = arc_a[3 * i], ay = arc_a[3 * i + 1], az = arc_a[3 * i + 2];
    const double ux = va[w] - va[u], uy = va[w + 1] - va[u + 1], uz = va[w + 2] - va[u + 2];
    const double ul = std::sqrt(ux * ux + uy * uy + uz * uz);
    for (int g = 0; g < FB; g++) side[g] = pb[4 * g] * ax + pb[4 * g + 1] * ay + pb[4 * g + 2] * az;
    for (int j = 0; j < EB; j++) {
      if (reach_b[j] == 0) continue;
      const double cba = side[sb[2 * j]], dba = side[sb[2 * j + 1]];
      if (!(cba * dba < 0)) continue;
      const double bx = arc_b[3 * j], by = arc_b[3 * j + 1], bz = arc_b[3 * j + 2];
      const double adc = pa[l] * bx + pa[l + 1] * by + pa[l + 2] * bz, bdc = pa[r] * bx + pa[r + 1] * by + pa[r + 2] * bz;
      if (!(adc * bdc < 0 && cba * bdc < 0)) continue;
      const int p = 3 * eb[2 * j], q = 3 * eb[2 * j + 1];
      const double tx = vb[q] - vb[p], ty = vb[q + 1] - vb[p + 1], tz = vb[q + 2] - vb[p + 2];
      const double len = normalize(S, uy * tz - uz * ty, uz * tx - ux * tz, ux * ty - uy * tx);
      if (!(len > PARALLEL * ul * std::sqrt(tx * tx + ty * ty + tz * tz))) continue;
      double nx = S.unit[0], ny = S.unit[1], nz = S.unit[2];
      if (nx * (va[u] - ca[0]) + ny * (va[u + 1] - ca[1]) + nz * (va[u + 2] - ca[2]) < 0) { nx = -nx; ny = -ny; nz = -nz; }
      const double sep = nx * (vb[p] - va[u]) + ny * (vb[p + 1] - va[u + 1]) + nz * (vb[p + 2] - va[u + 2]);
      if (sep > margin) return;
      if (sep > edge_sep) { edge_sep = sep; ei = i; ej = j; enx = nx; eny = ny; enz = nz; }
    }
  }
  const bool edge_best = ei >= 0 && edge_sep > face_a && edge_sep > face_b, b_best = !edge_best && face_b > face_a;
  if (ei >= 0 && edge_sep > js_max(face_a, face_b) + tolerance) {
    hull_edges(S, A, ei, B, ej, enx, eny, enz, margin, out);
  } else {
    const bool on_b = face_b > face_a + tolerance, sure = !edge_best && on_b == b_best;
    if (hull_face(S, on_b ? B : A, on_b ? fb : fa, on_b ? A : B, on_b, margin, tolerance, out, sure)) {
      const double behind = js_max(js_max(face_a, face_b), edge_sep) - (on_b ? face_b : face_a);
      if (behind > out.loss) out.loss = behind;
    } else if (!sure) {
      if (edge_best) hull_edges(S, A, ei, B, ej, enx, eny, enz, margin, out);
      else hull_face(S, b_best ? B : A, b_best ? fb : fa, b_best ? A : B, b_best, margin, tolerance, out, true);
    }
  }
  if (out.count == 0 && js_max(js_max(face_a, face_b), edge_sep) > 0) nearest(S, a, b, 0, margin, out);
}

// src/collide.js:1766-1781 (sphereHull)
void sphere_hull(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  Placed B;
  B.set(b);
  const double* p = B.planes.data();
  const double* c = a.center;
  const double r = a.half[0];
  double most = -INFINITY;
  int f = 0;
  for (int k = 0; k < B.hull->F; k++) {
    const double sep = p[4 * k] * c[0] + p[4 * k + 1] * c[1] + p[4 * k + 2] * c[2] - p[4 * k + 3];
    if (sep > most) { most = sep; f = k; }
  }
  if (most - r > margin) return;
  if (most > 0 && nearest(S, a, b, r, margin, out) > 0) return;
  const double nx = -p[4 * f], ny = -p[4 * f + 1], nz = -p[4 * f + 2];
  const double separation = most - r, t = r + separation * 0.5;
  set_normal(out, nx, ny, nz);
  add_point(out, c[0] + nx * t, c[1] + ny * t, c[2] + nz * t, separation, mix(4, f, 0, 0));
}

// src/collide.js:1855-1882 (clipToHullFace): the capsule's core (S.seg_a) clipped to face f's sides.
bool clip_to_hull_face(Scratch& S, const Placed& B, int f, double r, double margin, Manifold& out) {
  const double* p = B.planes.data();
  const Hull& h = *B.hull;
  const double* v = B.v.data();
  const double* sa = S.seg_a;
  const double nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
  const int i0 = h.start[f], k = h.start[f + 1] - i0;
  double u0 = 0, u1 = 1;
  for (int j = 0; j < k; j++) {
    const int sv = 3 * h.loops[i0 + j], q = 3 * h.loops[i0 + (j + 1 < k ? j + 1 : 0)];
    const double ex = v[q] - v[sv], ey = v[q + 1] - v[sv + 1], ez = v[q + 2] - v[sv + 2];
    const double mx = ey * nz - ez * ny, my = ez * nx - ex * nz, mz = ex * ny - ey * nx;
    const double offset = mx * v[sv] + my * v[sv + 1] + mz * v[sv + 2];
    const double d0 = sa[0] * mx + sa[1] * my + sa[2] * mz - offset, d1 = sa[3] * mx + sa[4] * my + sa[5] * mz - offset;
    if (d0 > 0 && d1 > 0) return false;
    if (d0 > 0) u0 = js_max(u0, d0 / (d0 - d1));
    else if (d1 > 0) u1 = js_min(u1, d0 / (d0 - d1));
  }
  if (u0 > u1) return false;
  set_normal(out, -nx, -ny, -nz);
  for (int e = 0; e < (u1 > u0 ? 2 : 1); e++) {
    const double u = e == 0 ? u0 : u1;
    const double x = sa[0] + (sa[3] - sa[0]) * u, y = sa[1] + (sa[4] - sa[1]) * u, z = sa[2] + (sa[5] - sa[2]) * u;
    const double above = x * nx + y * ny + z * nz - p[4 * f + 3], separation = above - r;
    if (separation > margin) continue;
    const double t = (above + r) * 0.5;
    add_point(out, x - nx * t, y - ny * t, z - nz * t, separation, mix(6, f, e, 0));
  }
  return out.count > 0;
}

// src/collide.js:1790-1847 (capsuleHull)
void capsule_hull(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  Placed B;
  B.set(b);
  const double* p = B.planes.data();
  const Hull& h = *B.hull;
  const int F = h.F;
  const double r = a.half[0];
  double* sa = S.seg_a;
  double* sb = S.seg_b;
  segment(a, sa);
  double na[3], nb[3];
  const double d = gjk_closest(a, b, na, nb);
  if (d > 0) {
    if (d - r > margin) return;
    const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
    int f = 0;
    double most = -INFINITY;
    for (int k = 0; k < F; k++) {
      const double dot = -(p[4 * k] * nx + p[4 * k + 1] * ny + p[4 * k + 2] * nz);
      if (dot > most) { most = dot; f = k; }
    }
    if (1 - most <= PARALLEL && clip_to_hull_face(S, B, f, r, margin, out)) return;
    nearest(S, a, b, r, margin, out);
    return;
  }
  const double dx = sa[3] - sa[0], dy = sa[4] - sa[1], dz = sa[5] - sa[2];
  const double dl = std::sqrt(dx * dx + dy * dy + dz * dz);
  double best_sep = -INFINITY, nx = 0, ny = 0, nz = 0;
  int best_face = -1, edge_i = -1;
  for (int k = 0; k < F; k++) {
    const double mx = p[4 * k], my = p[4 * k + 1], mz = p[4 * k + 2];
    const double sep = js_min(sa[0] * mx + sa[1] * my + sa[2] * mz, sa[3] * mx + sa[4] * my + sa[5] * mz) - p[4 * k + 3] - r;
    if (sep > best_sep) { best_sep = sep; best_face = k; }
  }
  const int* e = h.edges.data();
  const double* v = B.v.data();
  for (int i = 0; i < h.E; i++) {
    const int u = 3 * e[2 * i], w = 3 * e[2 * i + 1];
    const double tx = v[w] - v[u], ty = v[w + 1] - v[u + 1], tz = v[w + 2] - v[u + 2];
    const double l = normalize(S, dy * tz - dz * ty, dz * tx - dx * tz, dx * ty - dy * tx);
    if (!(l > PARALLEL * dl * std::sqrt(tx * tx + ty * ty + tz * tz))) continue;
    double mx = S.unit[0], my = S.unit[1], mz = S.unit[2];
    if (mx * (v[u] - B.center[0]) + my * (v[u + 1] - B.center[1]) + mz * (v[u + 2] - B.center[2]) < 0) { mx = -mx; my = -my; mz = -mz; }
    const double top = -B.lowest(-mx, -my, -mz);
    if (top - (mx * v[u] + my * v[u + 1] + mz * v[u + 2]) > tolerance) continue;
    const double sep = js_min(sa[0] * mx + sa[1] * my + sa[2] * mz, sa[3] * mx + sa[4] * my + sa[5] * mz) - r - top;
    if (sep > best_sep + tolerance) { best_sep = sep; edge_i = i; nx = mx; ny = my; nz = mz; }
  }
  if (best_sep > margin) return;
  if (edge_i < 0) {
    clip_to_hull_face(S, B, best_face, r, margin, out);
    return;
  }
  const int u = 3 * e[2 * edge_i], w = 3 * e[2 * edge_i + 1];
  for (int k = 0; k < 3; k++) { sb[k] = v[u + k]; sb[3 + k] = v[w + k]; }
  closest_segments(S, sa, sb);
  const double px = sa[0] + dx * S.seg_s, py = sa[1] + dy * S.seg_s, pz = sa[2] + dz * S.seg_s;
  const double qx = sb[0] + (sb[3] - sb[0]) * S.seg_t, qy = sb[1] + (sb[4] - sb[1]) * S.seg_t, qz = sb[2] + (sb[5] - sb[2]) * S.seg_t;
  const double separation = (px - qx) * nx + (py - qy) * ny + (pz - qz) * nz - r;
  set_normal(out, -nx, -ny, -nz);
  const double t = r + separation * 0.5;
  add_point(out, px - nx * t, py - ny * t, pz - nz * t, separation, mix(5, edge_i, 0, 0));
}

// src/collide.js:204-242 (collide, pair), for spheres and boxes.
int collide(const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  Scratch S;
  out.count = 0;
  out.loss = 0;
  S.shape_a = &a;
  S.shape_b = &b;
  const bool swap = a.kind > b.kind;
  const Shape& lo = swap ? b : a;
  const Shape& hi = swap ? a : b;
  if (lo.kind == SPHERE && hi.kind == SPHERE) sphere_sphere(S, lo, hi, margin, out);
  else if (lo.kind == SPHERE && hi.kind == CAPSULE) sphere_capsule(S, lo, hi, margin, out);
  else if (lo.kind == SPHERE && hi.kind == BOX) sphere_box(S, lo, hi, margin, out);
  else if (lo.kind == CAPSULE && hi.kind == CAPSULE) capsule_capsule(S, lo, hi, margin, tolerance, out);
  else if (lo.kind == CAPSULE && hi.kind == BOX) capsule_box(S, lo, hi, margin, tolerance, out);
  else if (lo.kind == BOX && hi.kind == BOX) box_box(S, lo, hi, margin, tolerance, out);
  else if (lo.kind == SPHERE && hi.kind == HULL) sphere_hull(S, lo, hi, margin, out);
  else if (lo.kind == CAPSULE && hi.kind == HULL) capsule_hull(S, lo, hi, margin, tolerance, out);
  else if ((lo.kind == BOX || lo.kind == HULL) && hi.kind == HULL) hull_hull(S, lo, hi, margin, tolerance, out);
  else return -1;
  if (swap) { out.normal[0] = -out.normal[0]; out.normal[1] = -out.normal[1]; out.normal[2] = -out.normal[2]; }
  return out.count;
}


// ---- distance between cores, src/gjk.js ----

constexpr double PROGRESS = 1.0 / 1073741824;  // 2^-30, src/gjk.js:47
constexpr int GJK_ROUNDS = 64;                  // src/gjk.js:50
constexpr double FLAT = 1.0 / 4503599627370496; // 2^-52, src/gjk.js:57
const uint8_t FACES[16] = {0, 1, 2, 3, 0, 1, 3, 2, 0, 2, 3, 1, 1, 2, 3, 0};

// src/gjk.js:66-68 (radius): what the core leaves out of a round shape.
double radius(const Shape& s) { return s.kind == SPHERE || s.kind == CAPSULE ? s.half[0] : 0; }

// src/gjk.js:70-107 (support) for a sphere's, a triangle's, a hull's, a capsule's or a box's core.
void gjk_support(const Shape& s, double x, double y, double z, double* out) {
  const double* c = s.center;
  const double* a = s.axes;
  const double* h = s.half;
  if (s.kind == SPHERE) {
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
  } else if (s.kind == TRIANGLE_KIND) {
    // The corner farthest along; the first of equals.
    const double* p = s.points;
    int best = 0;
    double most = -INFINITY;
    for (int i = 0; i < 3; i++) {
      const double d = p[3 * i] * x + p[3 * i + 1] * y + p[3 * i + 2] * z;
      if (d > most) { most = d; best = i; }
    }
    out[0] = p[3 * best]; out[1] = p[3 * best + 1]; out[2] = p[3 * best + 2];
  } else if (s.kind == HULL) {
    const std::vector<double>& v = s.hull->vertices;
    const int V = s.hull->V;
    const double lx = x * a[0] + y * a[1] + z * a[2], ly = x * a[3] + y * a[4] + z * a[5], lz = x * a[6] + y * a[7] + z * a[8];
    int best = 0;
    double most = -INFINITY;
    for (int i = 0; i < V; i++) {
      const double d = v[3 * i] * lx + v[3 * i + 1] * ly + v[3 * i + 2] * lz;
      if (d > most) { most = d; best = i; }
    }
    const double px = v[3 * best], py = v[3 * best + 1], pz = v[3 * best + 2];
    out[0] = c[0] + px * a[0] + py * a[3] + pz * a[6];
    out[1] = c[1] + px * a[1] + py * a[4] + pz * a[7];
    out[2] = c[2] + px * a[2] + py * a[5] + pz * a[8];
  } else if (s.kind == CAPSULE) {
    const double e = x * a[3] + y * a[4] + z * a[5] >= 0 ? h[1] : -h[1];
    out[0] = c[0] + e * a[3]; out[1] = c[1] + e * a[4]; out[2] = c[2] + e * a[5];
  } else {
    const double ex = x * a[0] + y * a[1] + z * a[2] >= 0 ? h[0] : -h[0];
    const double ey = x * a[3] + y * a[4] + z * a[5] >= 0 ? h[1] : -h[1];
    const double ez = x * a[6] + y * a[7] + z * a[8] >= 0 ? h[2] : -h[2];
    out[0] = c[0] + ex * a[0] + ey * a[3] + ez * a[6];
    out[1] = c[1] + ex * a[1] + ey * a[4] + ez * a[7];
    out[2] = c[2] + ex * a[2] + ey * a[5] + ez * a[8];
  }
}

// The simplex of src/gjk.js (W, PA, PB, L, count) and its reduction (172-277).
struct Simplex {
  double W[12], PA[12], PB[12], L[4];
  int count;

  void move(int to, int from) {
    if (to == from) return;
    for (int c = 0; c < 3; c++) {
      W[3 * to + c] = W[3 * from + c]; PA[3 * to + c] = PA[3 * from + c]; PB[3 * to + c] = PB[3 * from + c];
    }
  }
  void keep(int i, int j, int k, double li, double lj, double lk, int n) {
    move(0, i); L[0] = li;
    if (n > 1) { move(1, j); L[1] = lj; }
    if (n > 2) { move(2, k); L[2] = lk; }
    count = n;
  }
  void segment(int i, int j) {
    const double ax = W[3 * i], ay = W[3 * i + 1], az = W[3 * i + 2];
    const double ex = W[3 * j] - ax, ey = W[3 * j + 1] - ay, ez = W[3 * j + 2] - az;
    const double ee = ex * ex + ey * ey + ez * ez;
    const double t = ee > 0 ? -(ax * ex + ay * ey + az * ez) / ee : 0;
    if (t <= 0) keep(i, i, i, 1, 0, 0, 1);
    else if (t >= 1) keep(j, j, j, 1, 0, 0, 1);
    else keep(i, j, j, 1 - t, t, 0, 2);
  }
  // Ericson's closest point on a triangle to the origin, by its Voronoi regions.
  void triangle(int i, int j, int k) {
    const double ax = W[3 * i], ay = W[3 * i + 1], az = W[3 * i + 2];
    const double bx = W[3 * j], by = W[3 * j + 1], bz = W[3 * j + 2];
    const double cx = W[3 * k], cy = W[3 * k + 1], cz = W[3 * k + 2];
    const double abx = bx - ax, aby = by - ay, abz = bz - az, acx = cx - ax, acy = cy - ay, acz = cz - az;
    const double d1 = -(abx * ax + aby * ay + abz * az), d2 = -(acx * ax + acy * ay + acz * az);
    if (d1 <= 0 && d2 <= 0) return keep(i, i, i, 1, 0, 0, 1);
    const double d3 = -(abx * bx + aby * by + abz * bz), d4 = -(acx * bx + acy * by + acz * bz);
    if (d3 >= 0 && d4 <= d3) return keep(j, j, j, 1, 0, 0, 1);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      const double v = d1 / (d1 - d3);
      return keep(i, j, j, 1 - v, v, 0, 2);
    }
    const double d5 = -(abx * cx + aby * cy + abz * cz), d6 = -(acx * cx + acy * cy + acz * cz);
    if (d6 >= 0 && d5 <= d6) return keep(k, k, k, 1, 0, 0, 1);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
      const double w = d2 / (d2 - d6);
      return keep(i, k, k, 1 - w, w, 0, 2);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
      const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return keep(j, k, k, 1 - w, w, 0, 2);
    }
    const double sc = 1 / (va + vb + vc), v = vb * sc, w = vc * sc;
    keep(i, j, k, 1 - v - w, v, w, 3);
  }
  // The face of the tetrahedron nearest the origin, or false when the origin is inside.
  bool tetrahedron() {
    double best = INFINITY;
    int best_face = -1;
    double w[12], pa[12], pb[12];
    std::memcpy(w, W, sizeof w); std::memcpy(pa, PA, sizeof pa); std::memcpy(pb, PB, sizeof pb);
    bool outside = false;
    for (int f = 0; f < 4; f++) {
      const int i = FACES[4 * f], j = FACES[4 * f + 1], k = FACES[4 * f + 2], o = FACES[4 * f + 3];
      const double ax = w[3 * i], ay = w[3 * i + 1], az = w[3 * i + 2];
      const double ux = w[3 * j] - ax, uy = w[3 * j + 1] - ay, uz = w[3 * j + 2] - az;
      const double vx = w[3 * k] - ax, vy = w[3 * k + 1] - ay, vz = w[3 * k + 2] - az;
      const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      const double to_origin = -(nx * ax + ny * ay + nz * az);
      const double ox = w[3 * o] - ax, oy = w[3 * o + 1] - ay, oz = w[3 * o + 2] - az;
      const double to_fourth = nx * ox + ny * oy + nz * oz;
      const bool flat = to_fourth * to_fourth <= FLAT * (nx * nx + ny * ny + nz * nz) * (ox * ox + oy * oy + oz * oz);
      if (!flat && to_origin * to_fourth >= 0) continue;
      outside = true;
      std::memcpy(W, w, sizeof w); std::memcpy(PA, pa, sizeof pa); std::memcpy(PB, pb, sizeof pb);
      count = 4;
      triangle(i, j, k);
      double px = 0, py = 0, pz = 0;
      for (int m = 0; m < count; m++) {
        px += L[m] * W[3 * m]; py += L[m] * W[3 * m + 1]; pz += L[m] * W[3 * m + 2];
      }
      const double d = px * px + py * py + pz * pz;
      if (d < best) { best = d; best_face = f; }
    }
    if (!outside) return false;
    const int i = FACES[4 * best_face], j = FACES[4 * best_face + 1], k = FACES[4 * best_face + 2];
    std::memcpy(W, w, sizeof w); std::memcpy(PA, pa, sizeof pa); std::memcpy(PB, pb, sizeof pb);
    count = 4;
    triangle(i, j, k);
    return true;
  }
  bool reduce() {
    if (count == 2) { segment(0, 1); return true; }
    if (count == 3) { triangle(0, 1, 2); return true; }
    return tetrahedron();
  }
};

// src/gjk.js:114-155 (closest): the distance between two cores and the nearest points on each.
double gjk_closest(const Shape& a, const Shape& b, double* near_a, double* near_b) {
  Simplex S;
  double sa[3], sb[3];
  double vx = a.center[0] - b.center[0], vy = a.center[1] - b.center[1], vz = a.center[2] - b.center[2];
  if (vx == 0 && vy == 0 && vz == 0) vx = 1;
  gjk_support(a, -vx, -vy, -vz, sa);
  gjk_support(b, vx, vy, vz, sb);
  S.W[0] = sa[0] - sb[0]; S.W[1] = sa[1] - sb[1]; S.W[2] = sa[2] - sb[2];
  S.PA[0] = sa[0]; S.PA[1] = sa[1]; S.PA[2] = sa[2];
  S.PB[0] = sb[0]; S.PB[1] = sb[1]; S.PB[2] = sb[2];
  S.count = 1;
  S.L[0] = 1;
  vx = S.W[0]; vy = S.W[1]; vz = S.W[2];
  for (int round = 0; round < GJK_ROUNDS; round++) {
    const double vv = vx * vx + vy * vy + vz * vz;
    if (vv == 0) break;
    gjk_support(a, -vx, -vy, -vz, sa);
    gjk_support(b, vx, vy, vz, sb);
    const double wx = sa[0] - sb[0], wy = sa[1] - sb[1], wz = sa[2] - sb[2];
    if (vv - (vx * wx + vy * wy + vz * wz) <= PROGRESS * vv) break;
    bool repeat = false;
    for (int i = 0; i < S.count; i++)
      if (S.W[3 * i] == wx && S.W[3 * i + 1] == wy && S.W[3 * i + 2] == wz) repeat = true;
    if (repeat) break;
    const int i3 = 3 * S.count;
    S.W[i3] = wx; S.W[i3 + 1] = wy; S.W[i3 + 2] = wz;
    S.PA[i3] = sa[0]; S.PA[i3 + 1] = sa[1]; S.PA[i3 + 2] = sa[2];
    S.PB[i3] = sb[0]; S.PB[i3 + 1] = sb[1]; S.PB[i3 + 2] = sb[2];
    S.count++;
    if (!S.reduce()) { vx = vy = vz = 0; break; }
    vx = 0; vy = 0; vz = 0;
    for (int i = 0; i < S.count; i++) {
      vx += S.L[i] * S.W[3 * i]; vy += S.L[i] * S.W[3 * i + 1]; vz += S.L[i] * S.W[3 * i + 2];
    }
  }
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  for (int i = 0; i < S.count; i++) {
    ax += S.L[i] * S.PA[3 * i]; ay += S.L[i] * S.PA[3 * i + 1]; az += S.L[i] * S.PA[3 * i + 2];
    bx += S.L[i] * S.PB[3 * i]; by += S.L[i] * S.PB[3 * i + 1]; bz += S.L[i] * S.PB[3 * i + 2];
  }
  near_a[0] = ax; near_a[1] = ay; near_a[2] = az;
  near_b[0] = bx; near_b[1] = by; near_b[2] = bz;
  return std::sqrt(vx * vx + vy * vy + vz * vz);
}

// ---- level geometry: src/mesh.js, src/triangle.js ----

#include "level.h"

// The heightfields level records name (oracle_fields), by number.
std::vector<Heightfield> g_fields;

// ---- time of impact, src/impact.js ----

constexpr int IMPACT_ROUNDS = 32;   // src/impact.js:31
constexpr double UNSURE = -2;       // src/impact.js:34
constexpr double SPIN_CAP = 3.141592653589793;  // pi, src/impact.js:28

// What meet() reads of a body besides its shape: its pose, motion and gravity.
struct Motion {
  double q[4], v[3], w[3];
  double gravity_scale;  // 0 unless dynamic (src/contact.js:687)
  double reach;
  bool dynamic, capped;  // capped: dynamic, and not a plain sphere (src/impact.js:99)
};

// src/impact.js:91-111 (start) for both bodies: positions, velocities, the pull of
// gravity and spins, the spin cut to pi/dt for all but a plain sphere.
struct Flight {
  double P[6], V[6], G[6], WV[6], Q[8];
};
void start(Flight& F, const Shape& s, const Motion& m, int o, double dt) {
  for (int k = 0; k < 3; k++) {
    F.P[o + k] = s.center[k];
    F.V[o + k] = m.v[k];
    F.G[o + k] = s.weight[k];
    F.WV[o + k] = m.w[k];
  }
  if (m.capped) {
    const double most = SPIN_CAP / dt;
    const double w2 = F.WV[o] * F.WV[o] + F.WV[o + 1] * F.WV[o + 1] + F.WV[o + 2] * F.WV[o + 2];
    if (w2 > most * most) {
      const double k = most / std::sqrt(w2);
      F.WV[o] *= k; F.WV[o + 1] *= k; F.WV[o + 2] *= k;
    }
  }
  const int q4 = o == 0 ? 0 : 4;
  for (int k = 0; k < 4; k++) F.Q[q4 + k] = m.q[k];
}

// src/impact.js:141-149 (turned): (x, y, z, w) times q, normalized.
void turned(double x, double y, double z, double w, const double* q, double* out) {
  const double bx = q[0], by = q[1], bz = q[2], bw = q[3];
  const double rx = w * bx + x * bw + y * bz - z * by;
  const double ry = w * by + y * bw + z * bx - x * bz;
  const double rz = w * bz + z * bw + x * by - y * bx;
  const double rw = w * bw - x * bx - y * by - z * bz;
  const double l = 1 / std::sqrt(rx * rx + ry * ry + rz * rz + rw * rw);
  out[0] = rx * l; out[1] = ry * l; out[2] = rz * l; out[3] = rw * l;
}

// src/impact.js:125-139 (posed): where body o is at time t, falling and turning at its spin.
void posed(const Flight& F, int o, double t, double* c, double* q) {
  c[0] = F.P[o] + (F.V[o] + 0.5 * F.G[o] * t) * t;
  c[1] = F.P[o + 1] + (F.V[o + 1] + 0.5 * F.G[o + 1] * t) * t;
  c[2] = F.P[o + 2] + (F.V[o + 2] + 0.5 * F.G[o + 2] * t) * t;
  const int q4 = o == 0 ? 0 : 4;
  const double wx = F.WV[o], wy = F.WV[o + 1], wz = F.WV[o + 2], w = std::sqrt(wx * wx + wy * wy + wz * wz);
  if (w == 0 || t == 0) {
    for (int k = 0; k < 4; k++) q[k] = F.Q[q4 + k];
    return;
  }
  const double half = 0.5 * w * t, k = oracle::js_sin(half) / w;
  turned(wx * k, wy * k, wz * k, oracle::js_cos(half), F.Q + q4, q);
}

// src/collide.js:82-91 (Shape.place)
void place(Shape& s, const double* c, const double* q) {
  s.center[0] = c[0]; s.center[1] = c[1]; s.center[2] = c[2];
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  double* a = s.axes;
  a[0] = 1 - 2 * (y * y + z * z); a[1] = 2 * (x * y + z * w); a[2] = 2 * (x * z - y * w);
  a[3] = 2 * (x * y - z * w); a[4] = 1 - 2 * (x * x + z * z); a[5] = 2 * (y * z + x * w);
  a[6] = 2 * (x * z + y * w); a[7] = 2 * (y * z - x * w); a[8] = 1 - 2 * (x * x + y * y);
}

// src/impact.js:63-89 (impact): conservative advancement of the cores over [0, dt).
// t, -1 (they cannot meet within the step) or UNSURE; sa and sb are left posed at the last t tried.
double impact(Flight& F, Shape& sa, Shape& sb, const Motion& ma, const Motion& mb, double dt, double tolerance) {
  start(F, sa, ma, 0, dt);
  start(F, sb, mb, 3, dt);
  const double ra = radius(sa), rb = radius(sb);
  const double* WV = F.WV;
  const double spin = std::sqrt(WV[0] * WV[0] + WV[1] * WV[1] + WV[2] * WV[2]) * ma.reach
      + std::sqrt(WV[3] * WV[3] + WV[4] * WV[4] + WV[5] * WV[5]) * mb.reach;
  const double gx = F.G[3] - F.G[0], gy = F.G[4] - F.G[1], gz = F.G[5] - F.G[2];
  const double gl = std::sqrt(gx * gx + gy * gy + gz * gz);
  double t = 0;
  for (int round = 0; round < IMPACT_ROUNDS; round++) {
    double c[3], q[4];
    posed(F, 0, t, c, q);
    place(sa, c, q);
    posed(F, 3, t, c, q);
    place(sb, c, q);
    double na[3], nb[3];
    const double d = gjk_closest(sa, sb, na, nb), gap = d - ra - rb;
    if (gap <= tolerance) return t;
    const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
    const double vx = F.V[3] - F.V[0] + gx * t, vy = F.V[4] - F.V[1] + gy * t, vz = F.V[5] - F.V[2] + gz * t;
    const double rate = -(vx * nx + vy * ny + vz * nz) + gl * (dt - t) + spin;
    if (!(rate > 0)) return -1;
    t += gap / rate;
    if (t >= dt) return -1;
  }
  return UNSURE;
}

// src/contact.js:686-693 (moving): how far the pair may close within the step.
double moving(const Motion& ma, const Motion& mb, const double* g, double dt) {
  const double dx = mb.v[0] - ma.v[0], dy = mb.v[1] - ma.v[1], dz = mb.v[2] - ma.v[2];
  const double gs = mb.gravity_scale - ma.gravity_scale, gl = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) * std::fabs(gs);
  const double turning = std::sqrt(ma.w[0] * ma.w[0] + ma.w[1] * ma.w[1] + ma.w[2] * ma.w[2]) * ma.reach
      + std::sqrt(mb.w[0] * mb.w[0] + mb.w[1] * mb.w[1] + mb.w[2] * mb.w[2]) * mb.reach;
  return (std::sqrt(dx * dx + dy * dy + dz * dz) + 0.5 * gl * dt + turning) * dt;
}

double least_separation(const Manifold& m) {
  double s = INFINITY;
  for (int k = 0; k < m.count; k++)
    if (m.separations[k] < s) s = m.separations[k];
  return s;
}

// src/contact.js:740-761 (carried) for a whole body: a point held at body-relative
// (x, y, z) at the moment of impact, turned qt, to where it is at the step's start, turned q.
void carried(const Shape& s, const double* qt, const double* q, double x, double y, double z, double* out) {
  if (s.kind == SPHERE) { out[0] = x; out[1] = y; out[2] = z; return; }
  oracle::rotate(qt, x, y, z, true, out);
  if (s.kind != CAPSULE) {
    oracle::rotate(q, out[0], out[1], out[2], false, out);
    return;
  }
  const double half = s.half[1];
  const double c = out[1] < -half ? -half : out[1] > half ? half : out[1];
  oracle::rotate(qt, 0, c, 0, false, out);
  const double rx = x - out[0], ry = y - out[1], rz = z - out[2];
  oracle::rotate(q, 0, c, 0, false, out);
  out[0] += rx; out[1] += ry; out[2] += rz;
}

// src/contact.js:712-732 (back): a manifold found at the moment of impact, its points
// carried back with their bodies to the step's start; SHIFT keeps how far b's end is
// from a's there.
void back(const Flight& F, const Shape& sa, const Shape& sb, const Motion& ma, const Motion& mb, double t,
          Manifold& mf, double* shift) {
  double CA[3], QA[4], CB[3], QB[4];
  posed(F, 0, t, CA, QA);
  posed(F, 3, t, CB, QB);
  const double* pa = F.P;
  const double* pb = F.P + 3;
  const double nx = mf.normal[0], ny = mf.normal[1], nz = mf.normal[2];
  for (int k = 0; k < mf.count; k++) {
    const int k3 = 3 * k;
    const double half = 0.5 * mf.separations[k];
    const double px = mf.points[k3], py = mf.points[k3 + 1], pz = mf.points[k3 + 2];
    double v3[3];
    carried(sa, QA, ma.q, px - nx * half - CA[0], py - ny * half - CA[1], pz - nz * half - CA[2], v3);
    const double ax = pa[0] + v3[0], ay = pa[1] + v3[1], az = pa[2] + v3[2];
    carried(sb, QB, mb.q, px + nx * half - CB[0], py + ny * half - CB[1], pz + nz * half - CB[2], v3);
    const double bx = pb[0] + v3[0] - ax, by = pb[1] + v3[1] - ay, bz = pb[2] + v3[2] - az;
    mf.separations[k] = bx * nx + by * ny + bz * nz;
    mf.points[k3] = ax; mf.points[k3 + 1] = ay; mf.points[k3 + 2] = az;
    shift[k3] = bx; shift[k3 + 1] = by; shift[k3 + 2] = bz;
  }
}

// src/contact.js:627-640 (meet) for two whole bodies: the manifold now, or, when the pair
// is apart and closing faster than the tolerance, the one at the moment of impact carried
// back (shifted). Returns the point count (0: none), or -1 for a kind not restated here.
int meet(Shape sa, Shape sb, const Motion& ma, const Motion& mb, const double* gravity, double dt, double margin,
         double tolerance, Manifold& out, double* shift, bool& shifted, bool& asked) {
  shifted = false;
  asked = false;
  const int n = collide(sa, sb, margin, tolerance, out);
  if (n <= 0) return n;
  const double moves = least_separation(out) > tolerance ? moving(ma, mb, gravity, dt) : 0;
  if (moves > tolerance) {
    Flight F;
    const double t = impact(F, sa, sb, ma, mb, dt, tolerance);
    asked = true;
    const double ahead = t >= 0 ? moves * (1 - t / dt) + 2 * tolerance : 0;
    if (t == -1) return 0;
    if (t != UNSURE) {
      const int m = collide(sa, sb, ahead, tolerance, out);
      if (m <= 0) return m;
      back(F, sa, sb, ma, mb, t, out, shift);
      shifted = true;
    }
  }
  return out.count;
}

void read_shape(const double* r, Shape& s) {
  s.kind = (int)r[0];
  for (int i = 0; i < 3; i++) s.center[i] = r[1 + i];
  for (int i = 0; i < 9; i++) s.axes[i] = r[4 + i];
  for (int i = 0; i < 3; i++) s.half[i] = r[13 + i];
  for (int i = 0; i < 3; i++) s.weight[i] = r[16 + i];
  resolve(s);
}

// A body record for meet: the shape record (19), then q 4, v 3, w 3, gravity scale
// (0 unless dynamic), reach, flags (1 dynamic, 2 spin capped): 32 numbers.
void read_body(const double* r, Shape& s, Motion& m) {
  read_shape(r, s);
  for (int i = 0; i < 4; i++) m.q[i] = r[19 + i];
  for (int i = 0; i < 3; i++) m.v[i] = r[23 + i];
  for (int i = 0; i < 3; i++) m.w[i] = r[26 + i];
  m.gravity_scale = r[29];
  m.reach = r[30];
  const int flags = (int)r[31];
  m.dynamic = (flags & 1) != 0;
  m.capped = (flags & 2) != 0;
}

}  // namespace

extern "C" {

/* Shapes as flat records: kind, center 3, axes 9, half 3, weight 3 (19 numbers, kind as a
   double); results: count, normal 3, points 12, separations 4, ids 4 per pair. Returns the
   number of pairs done, or -(k+1) for pair k of a kind not restated here (capsules, hulls). */
int oracle_collide(int n, const double* a, const double* b, const double* margin, const double* tolerance,
                   int32_t* count, double* normal, double* points, double* separations, uint32_t* ids) {
  for (int k = 0; k < n; k++) {
    Shape sa, sb;
    Shape* sh[2] = {&sa, &sb};
    const double* src[2] = {a + 19 * k, b + 19 * k};
    for (int t = 0; t < 2; t++) {
      const double* r = src[t];
      sh[t]->kind = (int)r[0];
      for (int i = 0; i < 3; i++) sh[t]->center[i] = r[1 + i];
      for (int i = 0; i < 9; i++) sh[t]->axes[i] = r[4 + i];
      for (int i = 0; i < 3; i++) sh[t]->half[i] = r[13 + i];
      for (int i = 0; i < 3; i++) sh[t]->weight[i] = r[16 + i];
      resolve(*sh[t]);
    }
    Manifold m{};
    if (k == 0) g_loss.assign(n, 0);
    if (collide(sa, sb, margin[k], tolerance[k], m) < 0) return -(k + 1);
    g_loss[k] = m.loss;
    count[k] = m.count;
    for (int i = 0; i < 3; i++) normal[3 * k + i] = m.normal[i];
    for (int i = 0; i < 12; i++) points[12 * k + i] = m.points[i];
    for (int i = 0; i < 4; i++) separations[4 * k + i] = m.separations[i];
    for (int i = 0; i < 4; i++) ids[4 * k + i] = m.ids[i];
  }
  return n;
}

/* The hull table: n hulls, hull h's corners [hull_v[h], hull_v[h+1]) of `vertices` (3
   each), its faces [hull_f[h], hull_f[h+1]) of `planes` (4 each) with their loop starts in
   `start` (one a face, counted from the hull's first loop entry), its loops
   [hull_l[h], hull_l[h+1]) of `loops` (corner numbers of the hull). */
void oracle_hulls(int n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l, const double* vertices,
                  const int32_t* start, const int32_t* loops, const double* planes) {
  g_hulls.assign(n, Hull());
  for (int h = 0; h < n; h++) {
    Hull& x = g_hulls[h];
    x.vertices.assign(vertices + 3 * hull_v[h], vertices + 3 * hull_v[h + 1]);
    x.planes.assign(planes + 4 * hull_f[h], planes + 4 * hull_f[h + 1]);
    x.start.assign(start + hull_f[h], start + hull_f[h + 1]);
    x.start.push_back(hull_l[h + 1] - hull_l[h]);
    x.loops.assign(loops + hull_l[h], loops + hull_l[h + 1]);
    derive(x);
  }
}

/* The last oracle_collide's losses (src/collide.js Manifold.loss), one a pair. */
void oracle_losses(double* out) {
  for (size_t k = 0; k < g_loss.size(); k++) out[k] = g_loss[k];
}

/* src/gjk.js:114-155 on whole shapes (19-number records): distance and nearest points. */
void oracle_closest(int n, const double* a, const double* b, double* distance, double* near_a, double* near_b) {
  for (int k = 0; k < n; k++) {
    Shape sa, sb;
    read_shape(a + 19 * k, sa);
    read_shape(b + 19 * k, sb);
    distance[k] = gjk_closest(sa, sb, near_a + 3 * k, near_b + 3 * k);
  }
}

/* src/impact.js:63-89 on body records (32 numbers, see read_body); t, and the two
   shapes' centres and axes where it left them (12 numbers a shape). */
void oracle_impact(int n, const double* a, const double* b, const double* dt, const double* tolerance, double* t,
                   double* posed_a, double* posed_b) {
  for (int k = 0; k < n; k++) {
    Shape sa, sb;
    Motion ma, mb;
    read_body(a + 32 * k, sa, ma);
    read_body(b + 32 * k, sb, mb);
    Flight F;
    t[k] = impact(F, sa, sb, ma, mb, dt[k], tolerance[k]);
    for (int i = 0; i < 3; i++) { posed_a[12 * k + i] = sa.center[i]; posed_b[12 * k + i] = sb.center[i]; }
    for (int i = 0; i < 9; i++) { posed_a[12 * k + 3 + i] = sa.axes[i]; posed_b[12 * k + 3 + i] = sb.axes[i]; }
  }
}

/* src/contact.js:627-640 (meet) on body records (32 numbers) and per-pair scalars
   (margin, tolerance, dt, gravity 3: 6 numbers). Results as oracle_collide's, plus the
   shift (12 a pair) and flags (1 shifted, 2 a time of impact was asked). Returns n, or
   -(k+1) for pair k of a kind not restated here. */
int oracle_meet(int n, const double* a, const double* b, const double* pair, int32_t* count, double* normal,
                double* points, double* separations, uint32_t* ids, double* shift, int32_t* flags, double* loss) {
  for (int k = 0; k < n; k++) {
    Shape sa, sb;
    Motion ma, mb;
    read_body(a + 32 * k, sa, ma);
    read_body(b + 32 * k, sb, mb);
    const double* pr = pair + 6 * k;
    Manifold m{};
    double sh[12] = {0};
    bool shifted = false, asked = false;
    const int c = meet(sa, sb, ma, mb, pr + 3, pr[2], pr[0], pr[1], m, sh, shifted, asked);
    if (c < 0) return -(k + 1);
    count[k] = c;
    flags[k] = (shifted ? 1 : 0) | (asked ? 2 : 0);
    for (int i = 0; i < 3; i++) normal[3 * k + i] = m.normal[i];
    for (int i = 0; i < 12; i++) points[12 * k + i] = m.points[i];
    for (int i = 0; i < 4; i++) separations[4 * k + i] = m.separations[i];
    for (int i = 0; i < 4; i++) ids[4 * k + i] = m.ids[i];
    for (int i = 0; i < 12; i++) shift[12 * k + i] = sh[i];
    if (loss) loss[k] = m.loss;
  }
  return n;
}

/* src/core/math.js:102-122, for checking against the reference's own values. */
/* Heightfield tables: n fields, field f's heights heights[off[f] ..] (rows[f] rows of
   columns[f]), samples spacing[f] apart (src/mesh.js:288 Heightfield). */
void oracle_fields(int n, const int32_t* columns, const int32_t* rows, const double* spacing, const int64_t* off,
                   const double* heights) {
  g_fields.assign(n, Heightfield());
  for (int f = 0; f < n; f++) g_fields[f].init(heights + off[f], columns[f], rows[f], spacing[f]);
}

/* Field f's edge flags and folds for triangles [0, count): flags 1 a triangle, folds 6. */
void oracle_field_edges(int f, int count, int32_t* flags, double* folds) {
  for (int t = 0; t < count; t++) flags[t] = g_fields.at(f).edge(t, folds + 6 * t);
}

/* touchLevel (src/triangle.js:1078) for n shapes: shape k (a 19-number record) against
   field[k] placed at position 3k with axes 9k, looking in box 6k (the field's frame: lo 3,
   hi 3), margin and tolerance. Out: manifolds[k], then cap manifold slots a shape: count,
   normal 3, points 12, separations 4, ids 4. Returns n, or -(k+1) when shape k gives more
   than cap. */
int oracle_level(int n, const int32_t* field, const double* position, const double* axes, const double* shapes,
                 const double* boxes, const double* margin, const double* tolerance, int cap, int32_t* manifolds,
                 int32_t* count, double* normal, double* points, double* separations, uint32_t* ids) {
  LevelToucher L;
  std::vector<Manifold> out;
  for (int k = 0; k < n; k++) {
    Shape s;
    read_shape(shapes + 19 * k, s);
    const int got = L.touch_level(g_fields.at(field[k]), position + 3 * k, axes + 9 * k, s, boxes + 6 * k, margin[k],
                                  tolerance[k], out);
    if (got > cap) return -(k + 1);
    manifolds[k] = got;
    for (int j = 0; j < got; j++) {
      const size_t o = (size_t)k * cap + j;
      const Manifold& m = out[j];
      count[o] = m.count;
      for (int i = 0; i < 3; i++) normal[3 * o + i] = m.normal[i];
      for (int i = 0; i < 12; i++) points[12 * o + i] = i < 3 * m.count ? m.points[i] : 0;
      for (int i = 0; i < 4; i++) separations[4 * o + i] = i < m.count ? m.separations[i] : 0;
      for (int i = 0; i < 4; i++) ids[4 * o + i] = i < m.count ? m.ids[i] : 0;
    }
  }
  return n;
}

void oracle_sincos(int n, const double* x, double* s, double* c) {
  for (int k = 0; k < n; k++) {
    s[k] = oracle::js_sin(x[k]);
    c[k] = oracle::js_cos(x[k]);
  }
}

}  // extern "C"
