// The narrowphase on the MI355X for spheres, capsules and boxes, one pair a lane:
// src/collide.js collide / pair (204-242), the sphere and capsule pairs (349-676: closest
// points, the near-parallel overlap of two cores, a core clipped to a box face, a core
// through a box by separating axes) and box against box (713-1229:
// fifteen separating axes, the reference face clipped with snap slack, feature ids named
// by both boxes, the four-point reduction with its HELD UP / UNDER restarts, edge
// contact, closest points). Every branch and every operation is the reference's, in its
// order, in doubles without contraction: the manifolds are the reference's bits.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "collide.h"
#include "jsmath.h"

namespace bm {
namespace narrow {

__device__ __forceinline__ double js_min(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a < b) return a;
  if (b < a) return b;
  return __builtin_signbit(a) ? a : b;
}

constexpr int SPHERE = 0, CAPSULE = 1, BOX = 2, HULL = 3;
constexpr int MAX_POINTS = 4;
constexpr double PARALLEL = 7.450580596923828e-9;  // 2^-27, src/collide.js:686

struct Shape {
  int kind;
  double center[3];
  double axes[9];
  double half[3];
  double weight[3];
  int hull;                 // a hull's number in the table
  const HullTable* table;
  const double* points;     // a triangle's corners, in the world
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

__device__ constexpr int CORNER_P[4] = {-1, 1, 1, -1}, CORNER_Q[4] = {-1, -1, 1, 1};
__device__ constexpr int EDGE_FROM[12] = {0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3};
__device__ constexpr int EDGE_TO[12] = {1, 3, 5, 7, 2, 3, 6, 7, 4, 5, 6, 7};

// src/collide.js:860-877 (SWAPPED): a clipped point's feature named with the two boxes'
// roles swapped. Corners 0-3 of the incident face and corners 20 + 4j + k of the
// reference face name each other (27, 23, 22, 26 for corners 0-3); an incident edge e
// cut by side plane k, 4 + 4e + k, is the other box's edge on plane k cut by the plane
// edge e lies on (EDGE_PLANE {3, 0, 2, 1}, PLANE_EDGE {1, 3, 2, 0}); the rest stay.
__device__ __forceinline__ uint32_t swapped(int f) {
  if (f < 4) return (0x1A16171Bu >> (8 * f)) & 0xffu;
  if (f < 20) {
    const int e = (f - 4) >> 2, k = (f - 4) & 3;
    return 4 + 4 * ((0x0231u >> (4 * k)) & 0xfu) + ((0x1203u >> (4 * e)) & 0xfu);
  }
  if (f == 27) return 0;
  if (f == 23) return 1;
  if (f == 22) return 2;
  if (f == 26) return 3;
  return (uint32_t)f;
}

__device__ __forceinline__ double sign(double x) { return x < 0 ? -1 : 1; }
__device__ __forceinline__ double clamp01(double x) { return x < 0 ? 0 : x > 1 ? 1 : x; }

// src/collide.js:262-270
__device__ __forceinline__ double normalize(Scratch& S, double x, double y, double z) {
  const double m = js_max(js_max(dabs(x), dabs(y)), dabs(z));
  if (m == 0) return 0;
  const double sx = x / m, sy = y / m, sz = z / m;
  const double l = dsqrt(sx * sx + sy * sy + sz * sz);
  S.unit[0] = sx / l; S.unit[1] = sy / l; S.unit[2] = sz / l;
  return l * m;
}

__device__ __forceinline__ void set_normal(Manifold& out, double x, double y, double z) {
  out.normal[0] = x; out.normal[1] = y; out.normal[2] = z;
}

__device__ __forceinline__ void add_point(Manifold& out, double x, double y, double z, double separation, uint32_t id) {
  const int k = out.count++;
  out.points[3 * k] = x; out.points[3 * k + 1] = y; out.points[3 * k + 2] = z;
  out.separations[k] = separation;
  out.ids[k] = id;
}

// src/collide.js:297-306
__device__ __forceinline__ bool balls(Scratch& S, double ax, double ay, double az, double ra, double bx, double by, double bz, double rb,
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
__device__ __forceinline__ void closest_segments(Scratch& S, const double* p, const double* q) {
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
__device__ __forceinline__ void sphere_sphere(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  balls(S, a.center[0], a.center[1], a.center[2], a.half[0], b.center[0], b.center[1], b.center[2], b.half[0], 0, 1, 0,
        margin, out, 0);
}

// src/collide.js:367-403
__device__ __forceinline__ void sphere_box(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  const double* ax = b.axes;
  const double* h = b.half;
  const double r = a.half[0];
  const double dx = a.center[0] - b.center[0], dy = a.center[1] - b.center[1], dz = a.center[2] - b.center[2];
  const double lx = dx * ax[0] + dy * ax[1] + dz * ax[2];
  const double ly = dx * ax[3] + dy * ax[4] + dz * ax[5];
  const double lz = dx * ax[6] + dy * ax[7] + dz * ax[8];
  const bool inside = dabs(lx) <= h[0] && dabs(ly) <= h[1] && dabs(lz) <= h[2];
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
  double depth = h[0] - dabs(lx);
  if (h[1] - dabs(ly) < depth) { axis = 1; depth = h[1] - dabs(ly); }
  if (h[2] - dabs(lz) < depth) { axis = 2; depth = h[2] - dabs(lz); }
  const double l = axis == 0 ? lx : axis == 1 ? ly : lz;
  const double s = -sign(l);
  const double nx = ax[3 * axis] * s, ny = ax[3 * axis + 1] * s, nz = ax[3 * axis + 2] * s;
  set_normal(out, nx, ny, nz);
  const double separation = -(depth + r);
  const double t = r + separation * 0.5;
  add_point(out, a.center[0] + nx * t, a.center[1] + ny * t, a.center[2] + nz * t, separation, 0);
}

// src/collide.js:271-281 (perpendicular): a unit vector at right angles to v, into unit.
__device__ __forceinline__ void perpendicular(Scratch& S, double vx, double vy, double vz) {
  const double ax = dabs(vx), ay = dabs(vy), az = dabs(vz);
  if (ax <= ay && ax <= az) normalize(S, 0, vz, -vy);
  else if (ay <= az) normalize(S, -vz, 0, vx);
  else normalize(S, vy, -vx, 0);
}

// src/collide.js:308-312 (segment): a capsule's core segment.
__device__ __forceinline__ void segment(const Shape& c, double* out) {
  const double h = c.half[1];
  const double* y = c.axes;
  out[0] = c.center[0] - y[3] * h; out[1] = c.center[1] - y[4] * h; out[2] = c.center[2] - y[5] * h;
  out[3] = c.center[0] + y[3] * h; out[4] = c.center[1] + y[4] * h; out[5] = c.center[2] + y[5] * h;
}

// src/collide.js:314-319 (nearestOnSegment)
__device__ __forceinline__ double nearest_on_segment(const double* s, double x, double y, double z) {
  const double dx = s[3] - s[0], dy = s[4] - s[1], dz = s[5] - s[2];
  const double dd = dx * dx + dy * dy + dz * dz;
  if (dd == 0) return 0;
  return clamp01(((x - s[0]) * dx + (y - s[1]) * dy + (z - s[2]) * dz) / dd);
}

// src/collide.js:357-364 (sphereCapsule)
__device__ __forceinline__ void sphere_capsule(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
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
__device__ __forceinline__ void capsule_capsule(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
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
      double deepest = __builtin_inf();
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

__device__ __forceinline__ void to_world(Scratch& S, const Shape& b, double x, double y, double z);

// src/collide.js:559-575 (segmentHitsBox): whether the segment (box frame) passes through the box.
__device__ __forceinline__ bool segment_hits_box(const double* s, const double* h) {
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
__device__ __forceinline__ bool clip_to_face(Scratch& S, const Shape& b, int axis, double facing, double r, double margin, Manifold& out,
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
__device__ __forceinline__ void capsule_box_deep(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  const double* h = b.half;
  const double r = a.half[0];
  const double* local = S.local;
  double* edge = S.edge;
  const double dx = local[3] - local[0], dy = local[4] - local[1], dz = local[5] - local[2];
  double best_sep = -__builtin_inf(), nx = 0, ny = 0, nz = 0;
  int best_axis = -1;
  for (int c = 0; c < 6; c++) {
    double x = 0, y = 0, z = 0;
    if (c < 3) {
      if (c == 0) x = 1; else if (c == 1) y = 1; else z = 1;
    } else {
      const int e = c - 3;
      const double l = normalize(S, e == 0 ? 0 : e == 1 ? -dz : dy, e == 0 ? dz : e == 1 ? 0 : -dx,
                                 e == 0 ? -dy : e == 1 ? dx : 0);
      if (!(l > PARALLEL * dsqrt(dx * dx + dy * dy + dz * dz))) continue;
      x = S.unit[0]; y = S.unit[1]; z = S.unit[2];
    }
    const double sup = h[0] * dabs(x) + h[1] * dabs(y) + h[2] * dabs(z);
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
__device__ __forceinline__ void capsule_box(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
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
  double best = __builtin_inf(), sx = 0, sy = 0, sz = 0, qx = 0, qy = 0, qz = 0;
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
  const bool on_x = dabs(qx) == h[0], on_y = dabs(qy) == h[1], on_z = dabs(qz) == h[2];
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
__device__ __forceinline__ double support(const double* X, const double* h, double nx, double ny, double nz) {
  return h[0] * dabs(X[0] * nx + X[1] * ny + X[2] * nz) + h[1] * dabs(X[3] * nx + X[4] * ny + X[5] * nz)
      + h[2] * dabs(X[6] * nx + X[7] * ny + X[8] * nz);
}

// src/collide.js:552-557
__device__ __forceinline__ void to_world(Scratch& S, const Shape& b, double x, double y, double z) {
  const double* ax = b.axes;
  S.world[0] = b.center[0] + ax[0] * x + ax[3] * y + ax[6] * z;
  S.world[1] = b.center[1] + ax[1] * x + ax[4] * y + ax[7] * z;
  S.world[2] = b.center[2] + ax[2] * x + ax[5] * y + ax[8] * z;
}

// src/collide.js:780-789
__device__ __forceinline__ void corners(const Shape& x, double* c) {
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
__device__ __forceinline__ void box_box_closest(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  corners(a, S.corner_a);
  corners(b, S.corner_b);
  double best = __builtin_inf(), ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
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
__device__ __forceinline__ int clip_polygon(const double* src, const int* sf, const int* sl, int count, double* dst, int* df, int* dl, double px,
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
__device__ __forceinline__ void four(Scratch& S, int n, int a, double nx, double ny, double nz, double pressed, double touching) {
  const double* kept = S.rk;
  const double* sep = S.rs;
  int b = -1;
  double far = -1;
  for (int pass = 0; pass < 3 && b < 0; pass++) {
    const double within = pass == 0 ? pressed : pass == 1 ? touching : __builtin_inf();
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
    const double within = pass == 0 ? pressed : pass == 1 ? touching : __builtin_inf();
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
__device__ __forceinline__ bool foot(Scratch& S, const Shape& s, double nx, double ny, double nz) {
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
__device__ __forceinline__ bool rests(Scratch& S, const Shape& s, double nx, double ny, double nz) {
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
__device__ __forceinline__ int under(Scratch& S, int n, double nx, double ny, double nz, double pressed) {
  int start = -1;
  double near = __builtin_inf();
  for (int k = 0; k < n; k++) {
    if (S.rs[k] > pressed) continue;
    const double x = S.rk[3 * k] - S.foot[0], y = S.rk[3 * k + 1] - S.foot[1], z = S.rk[3 * k + 2] - S.foot[2];
    const double along = x * nx + y * ny + z * nz;
    const double d = x * x + y * y + z * z - along * along;
    if (d < near) { near = d; start = k; }
  }
  return start;
}

// src/collide.js:1077-1104 (reducePoints)
__device__ __forceinline__ void reduce_list(Scratch& S, const double* kept, const double* sep, const uint32_t* id, int n, double nx, double ny,
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
__device__ __forceinline__ void reduce_points(Scratch& S, int n, double nx, double ny, double nz, double tolerance,
                                              Manifold& out) {
  reduce_list(S, S.kept, S.kept_sep, S.kept_id, n, nx, ny, nz, tolerance, out);
}

// src/collide.js:884-945 (faceContact)
__device__ __forceinline__ void face_contact(Scratch& S, const Shape& ref, const Shape& inc, int axis, bool flipped, double margin,
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
    const double d = dabs(I[3 * j] * nx + I[3 * j + 1] * ny + I[3 * j + 2] * nz);
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
    S.kept_id[n] = flipped ? (inc_face << 9) | (ref_face << 6) | swapped(feat[v])
                           : (ref_face << 9) | (inc_face << 6) | (uint32_t)feat[v];
    n++;
  }
  if (n == 0) return;
  if (flipped) set_normal(out, -nx, -ny, -nz);
  else set_normal(out, nx, ny, nz);
  reduce_points(S, n, nx, ny, nz, tolerance, out);
}

// src/collide.js:1232-1243 (boxEdge)
__device__ __forceinline__ void box_edge(const Shape& x, int i, double nx, double ny, double nz, double* s) {
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
__device__ __forceinline__ void edge_contact(Scratch& S, const Shape& a, const Shape& b, int i, int j, double nx, double ny, double nz,
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
__device__ __forceinline__ void box_box(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  const double* A = a.axes;
  const double* B = b.axes;
  const double* ha = a.half;
  const double* hb = b.half;
  const double dx = b.center[0] - a.center[0], dy = b.center[1] - a.center[1], dz = b.center[2] - a.center[2];
  double* R = S.cosines;
  double face_a = -__builtin_inf(), face_b = -__builtin_inf();
  int axis_a = 0, axis_b = 0;
  for (int i = 0; i < 3; i++) {
    const double nx = A[3 * i], ny = A[3 * i + 1], nz = A[3 * i + 2];
    const double r0 = dabs(B[0] * nx + B[1] * ny + B[2] * nz);
    const double r1 = dabs(B[3] * nx + B[4] * ny + B[5] * nz);
    const double r2 = dabs(B[6] * nx + B[7] * ny + B[8] * nz);
    R[3 * i] = r0; R[3 * i + 1] = r1; R[3 * i + 2] = r2;
    const double s = dabs(dx * nx + dy * ny + dz * nz) - ha[i] - (hb[0] * r0 + hb[1] * r1 + hb[2] * r2);
    if (s > margin) return;
    if (s > face_a) { face_a = s; axis_a = i; }
  }
  for (int j = 0; j < 3; j++) {
    const double nx = B[3 * j], ny = B[3 * j + 1], nz = B[3 * j + 2];
    const double s = dabs(dx * nx + dy * ny + dz * nz) - hb[j] - (ha[0] * R[j] + ha[1] * R[3 + j] + ha[2] * R[6 + j]);
    if (s > margin) return;
    if (s > face_b) { face_b = s; axis_b = j; }
  }
  double edge_sep = -__builtin_inf(), enx = 0, eny = 0, enz = 0;
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
      const double s = dabs(d) - support(A, ha, nx, ny, nz) - support(B, hb, nx, ny, nz);
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

// ---- hulls: src/collide.js:1262-1882 ----

constexpr double LOOSE = 1.0 / 65536;      // 2^-16, src/collide.js:1270
constexpr double NEARLY = 1.0 / 67108864;  // 2^-26, src/collide.js:1278
constexpr int CUT = 1 << 20;               // src/collide.js:1541

// src/collide.js:1286-1295 (reaches)
__device__ __forceinline__ bool reaches(double a1, double a2, double c, double bound) {
  const double hi = a1 > a2 ? a1 : a2, lo = a1 > a2 ? a2 : a1;
  if (!(hi < bound)) return true;
  const double s = 1 - c * c;
  if (!(s >= NEARLY)) return true;
  const double u = lo - c * hi;
  return u >= 0 && !(dsqrt(hi * hi + u * u / s) < bound);
}

// src/collide.js:1309-1388 (Placed): a hull, or a box as the unit cube scaled, placed in the
// world. Corners, planes and edge arcs are worked out where they are read, each the same
// operations in the same order as the reference's set(), so every read has set()'s bits
// and nothing is kept in per-lane memory.
struct Placed {
  const double *hv, *hp;
  const int *start, *loops, *edges, *sides;
  int V, F, E;
  bool box;
  double c[3], A[9], half[3];
  double inner, reach;

  __device__ __forceinline__ void set(const Shape& s) {
    const HullTable& T = *s.table;
    box = s.kind == BOX;
    const int h = box ? T.cube : s.hull;
    const int* info = T.info + 8 * h;
    hv = T.vertices + 3 * info[0];
    hp = T.planes + 4 * info[1];
    start = T.start + info[2];
    loops = T.loops + info[3];
    edges = T.edges + 2 * info[4];
    sides = T.sides + 2 * info[4];
    V = info[5]; F = info[6]; E = info[7];
    for (int k = 0; k < 3; k++) { c[k] = s.center[k]; half[k] = s.half[k]; }
    for (int k = 0; k < 9; k++) A[k] = s.axes[k];
    inner = box ? js_min(js_min(half[0], half[1]), half[2]) : T.dinfo[2 * h];
    reach = box ? dsqrt(half[0] * half[0] + half[1] * half[1] + half[2] * half[2]) : T.dinfo[2 * h + 1];
  }
  __device__ __forceinline__ void vert(int i, double* o) const {
    double x = hv[3 * i], y = hv[3 * i + 1], z = hv[3 * i + 2];
    if (box) { x *= 2 * half[0]; y *= 2 * half[1]; z *= 2 * half[2]; }
    o[0] = c[0] + A[0] * x + A[3] * y + A[6] * z;
    o[1] = c[1] + A[1] * x + A[4] * y + A[7] * z;
    o[2] = c[2] + A[2] * x + A[5] * y + A[8] * z;
  }
  __device__ __forceinline__ void plane(int f, double* o) const {
    const double x = hp[4 * f], y = hp[4 * f + 1], z = hp[4 * f + 2];
    const double nx = A[0] * x + A[3] * y + A[6] * z, ny = A[1] * x + A[4] * y + A[7] * z, nz = A[2] * x + A[5] * y + A[8] * z;
    const double d = box ? dabs(x) * half[0] + dabs(y) * half[1] + dabs(z) * half[2] : hp[4 * f + 3];
    o[0] = nx; o[1] = ny; o[2] = nz;
    o[3] = d + nx * c[0] + ny * c[1] + nz * c[2];
  }
  __device__ __forceinline__ void arc(int e, double* o) const {
    double l[4], r[4];
    plane(sides[2 * e], l);
    plane(sides[2 * e + 1], r);
    o[0] = r[1] * l[2] - r[2] * l[1];
    o[1] = r[2] * l[0] - r[0] * l[2];
    o[2] = r[0] * l[1] - r[1] * l[0];
  }
  __device__ __forceinline__ double lowest(double nx, double ny, double nz) const {
    double m = __builtin_inf();
    for (int i = 0; i < V; i++) {
      double v[3];
      vert(i, v);
      const double d = v[0] * nx + v[1] * ny + v[2] * nz;
      if (d < m) m = d;
    }
    return m;
  }
};

// src/collide.js:1401-1404 (mix)
__device__ __forceinline__ uint32_t mix(uint32_t kind, uint32_t i, uint32_t j, uint32_t k) {
  return ((kind + 1) * 0x9e3779b1u) ^ ((i + 1) * 0x85ebca77u) ^ ((j + 1) * 0xc2b2ae3du) ^ ((k + 1) * 0x27d4eb2fu);
}

__device__ __forceinline__ double gjk_closest(const Shape& a, const Shape& b, double* near_a, double* near_b);

// src/collide.js:1751-1759 (nearest)
__device__ __forceinline__ double nearest(Scratch& S, const Shape& a, const Shape& b, double ra, double margin,
                                          Manifold& out) {
  double na[3], nb[3];
  const double d = gjk_closest(a, b, na, nb);
  if (!(d > 0) || d - ra > margin) return d;
  const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
  const double separation = d - ra, t = ra + separation * 0.5;
  set_normal(out, nx, ny, nz);
  add_point(out, na[0] + nx * t, na[1] + ny * t, na[2] + nz * t, separation, mix(3, 0, 0, 0));
  return d;
}

// src/collide.js:1570-1581 (named)
__device__ __forceinline__ uint32_t named(int f, bool flipped, int rf, int inc) {
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

// src/collide.js:1698-1725 (clipSides)
__device__ __forceinline__ int clip_sides(const double* src, const int* sf, const int* sl, int count, double* dst,
                                          int* df, int* dl, double mx, double my, double mz, double offset, int plane,
                                          double slack, double length, int cap) {
  int n = 0;
  // A convex polygon cut by a plane gains at most one corner, so `cap` (the reference sizes
  // its arrays the same) is never reached; the check keeps every write inside the arrays.
  for (int v = 0; v < count && n + 2 <= cap; v++) {
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

constexpr int CLIP_MAX = 2 * HULL_MAX_CORNERS + 2;

// src/collide.js:1592-1689 (hullFace)
__device__ __forceinline__ bool hull_face(Scratch& S, const Placed& R, int rf, const Placed& I, bool flipped,
                                          double margin, double tolerance, Manifold& out, bool sure) {
  double rpf[4];
  R.plane(rf, rpf);
  const double nx = rpf[0], ny = rpf[1], nz = rpf[2], d = rpf[3];
  int low = 0;
  double lowest = __builtin_inf();
  for (int i = 0; i < I.V; i++) {
    double v[3];
    I.vert(i, v);
    const double dot = v[0] * nx + v[1] * ny + v[2] * nz;
    if (dot < lowest) { lowest = dot; low = i; }
  }
  int inc = 0;
  double least = __builtin_inf();
  for (int f = 0; f < I.F; f++) {
    double pf[4];
    I.plane(f, pf);
    const double dot = pf[0] * nx + pf[1] * ny + pf[2] * nz;
    if (!(dot < least)) continue;
    bool at = false;
    for (int j = I.start[f]; j < I.start[f + 1] && !at; j++) at = I.loops[j] == low;
    if (at) { least = dot; inc = f; }
  }
  const int i0 = I.start[inc], m = I.start[inc + 1] - i0, r0 = R.start[rf], k = R.start[rf + 1] - r0;
  double clip_p[3 * CLIP_MAX], clip_q[3 * CLIP_MAX], kept_p[3 * (CLIP_MAX + 1)], kept_s[CLIP_MAX + 1];
  int feat_p[CLIP_MAX], feat_q[CLIP_MAX], line_p[CLIP_MAX], line_q[CLIP_MAX];
  uint32_t kept_i[CLIP_MAX + 1];
  for (int j = 0; j < m; j++) {
    I.vert(I.loops[i0 + j], &clip_p[3 * j]);
    feat_p[j] = j; line_p[j] = j;
  }
  int count = m, src = 0;
  for (int j = 0; j < k && count > 0; j++) {
    double rp[3], rq[3];
    R.vert(R.loops[r0 + j], rp);
    R.vert(R.loops[r0 + (j + 1 < k ? j + 1 : 0)], rq);
    const double ex = rq[0] - rp[0], ey = rq[1] - rp[1], ez = rq[2] - rp[2];
    const double mx = ey * nz - ez * ny, my = ez * nx - ex * nz, mz = ex * ny - ey * nx;
    const double offset = mx * rp[0] + my * rp[1] + mz * rp[2];
    const double length = dsqrt(mx * mx + my * my + mz * mz);
    count = src == 0 ? clip_sides(clip_p, feat_p, line_p, count, clip_q, feat_q, line_q, mx, my, mz, offset, j,
                                  tolerance * length, length, CLIP_MAX)
                     : clip_sides(clip_q, feat_q, line_q, count, clip_p, feat_p, line_p, mx, my, mz, offset, j,
                                  tolerance * length, length, CLIP_MAX);
    src = 1 - src;
  }
  const double* poly = src == 0 ? clip_p : clip_q;
  const int* feat = src == 0 ? feat_p : feat_q;
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
    double deepest = __builtin_inf();
    for (int j = 0; j < n; j++)
      if (kept_s[j] < deepest) deepest = kept_s[j];
    if (deepest > s + tolerance) {
      if (!sure) return false;
      int j = 0;
      while (I.loops[i0 + j] != low) j++;
      double v[3];
      I.vert(low, v);
      const double h = s * 0.5;
      kept_p[3 * n] = v[0] - nx * h; kept_p[3 * n + 1] = v[1] - ny * h; kept_p[3 * n + 2] = v[2] - nz * h;
      kept_s[n] = s;
      kept_i[n] = named(j, flipped, rf, inc);
      n++;
    }
  }
  if (n == 0) return false;
  if (flipped) set_normal(out, -nx, -ny, -nz);
  else set_normal(out, nx, ny, nz);
  reduce_list(S, kept_p, kept_s, kept_i, n, nx, ny, nz, tolerance, out);
  if (s <= tolerance) {
    double kept = __builtin_inf();
    for (int j = 0; j < out.count; j++)
      if (out.separations[j] < kept) kept = out.separations[j];
    if (kept - s > out.loss) out.loss = kept - s;
  }
  return true;
}

// src/collide.js:1728-1744 (hullEdges)
__device__ __forceinline__ void hull_edges(Scratch& S, const Placed& A, int i, const Placed& B, int j, double nx,
                                           double ny, double nz, double margin, Manifold& out) {
  double* sa = S.seg_a;
  double* sb = S.seg_b;
  A.vert(A.edges[2 * i], sa);
  A.vert(A.edges[2 * i + 1], sa + 3);
  B.vert(B.edges[2 * j], sb);
  B.vert(B.edges[2 * j + 1], sb + 3);
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

// src/collide.js:1406-1536 (hullHull)
__device__ __forceinline__ void hull_hull(Scratch& S, const Shape& a, const Shape& b, double margin, double tolerance,
                                          Manifold& out) {
  Placed A, B;
  A.set(a);
  B.set(b);
  const double* ca = A.c;
  const double* cb = B.c;
  const int FA = A.F, FB = B.F;
  const double loose = LOOSE * (dabs(ca[0]) + dabs(ca[1]) + dabs(ca[2]) + dabs(cb[0]) + dabs(cb[1]) + dabs(cb[2])
                                + A.reach + B.reach);
  const double ra = A.inner > loose ? A.inner - loose : -__builtin_inf();
  const double rb = B.inner > loose ? B.inner - loose : -__builtin_inf();
  double above_a[HULL_MAX_FACES], above_b[HULL_MAX_FACES];
  int top = 0;
  for (int f = 0; f < FA; f++) {
    double pf[4];
    A.plane(f, pf);
    above_a[f] = pf[0] * cb[0] + pf[1] * cb[1] + pf[2] * cb[2] - pf[3];
    if (above_a[f] > above_a[top]) top = f;
  }
  double pt[4];
  A.plane(top, pt);
  const double first_a = B.lowest(pt[0], pt[1], pt[2]) - pt[3];
  if (first_a > margin) return;
  double face_a = -__builtin_inf();
  int fa = 0;
  for (int f = 0; f < FA; f++) {
    double sep = first_a;
    if (f != top) {
      if (above_a[f] - rb < js_max(face_a, first_a)) continue;
      double pf[4];
      A.plane(f, pf);
      sep = B.lowest(pf[0], pf[1], pf[2]) - pf[3];
      if (sep > margin) return;
    }
    if (sep > face_a) { face_a = sep; fa = f; }
  }
  top = 0;
  for (int f = 0; f < FB; f++) {
    double pf[4];
    B.plane(f, pf);
    above_b[f] = pf[0] * ca[0] + pf[1] * ca[1] + pf[2] * ca[2] - pf[3];
    if (above_b[f] > above_b[top]) top = f;
  }
  B.plane(top, pt);
  const double first_b = A.lowest(pt[0], pt[1], pt[2]) - pt[3];
  if (first_b > margin) return;
  double face_b = -__builtin_inf();
  int fb = 0;
  for (int f = 0; f < FB; f++) {
    double sep = first_b;
    if (f != top) {
      if (above_b[f] - ra < js_max(js_max(face_b, first_b), face_a)) continue;
      double pf[4];
      B.plane(f, pf);
      sep = A.lowest(pf[0], pf[1], pf[2]) - pf[3];
      if (sep > margin) return;
    }
    if (sep > face_b) { face_b = sep; fb = f; }
  }
  const int EA = A.E, EB = B.E;
  double side[HULL_MAX_FACES];
  const double best = js_max(face_a, face_b);
  bool reach_b[HULL_MAX_EDGES];
  int reachable = 0;
  for (int j = 0; j < EB; j++) {
    double l[4], r[4], p[3];
    B.plane(B.sides[2 * j], l);
    B.plane(B.sides[2 * j + 1], r);
    B.vert(B.edges[2 * j], p);
    const double dx = ca[0] - p[0], dy = ca[1] - p[1], dz = ca[2] - p[2];
    const bool can = reaches(l[0] * dx + l[1] * dy + l[2] * dz, r[0] * dx + r[1] * dy + r[2] * dz,
                             l[0] * r[0] + l[1] * r[1] + l[2] * r[2], best + ra);
    reach_b[j] = can;
    if (can) reachable++;
  }
  double edge_sep = -__builtin_inf(), enx = 0, eny = 0, enz = 0;
  int ei = -1, ej = -1;
  for (int i = 0; i < EA && reachable > 0; i++) {
    double l[4], r[4], u[3], w[3];
    A.plane(A.sides[2 * i], l);
    A.plane(A.sides[2 * i + 1], r);
    A.vert(A.edges[2 * i], u);
    A.vert(A.edges[2 * i + 1], w);
    const double dx = cb[0] - u[0], dy = cb[1] - u[1], dz = cb[2] - u[2];
    if (!reaches(l[0] * dx + l[1] * dy + l[2] * dz, r[0] * dx + r[1] * dy + r[2] * dz,
                 l[0] * r[0] + l[1] * r[1] + l[2] * r[2], best + rb))
      continue;
    double ar[3];
    A.arc(i, ar);
    const double ax = ar[0], ay = ar[1], az = ar[2];
    const double ux = w[0] - u[0], uy = w[1] - u[1], uz = w[2] - u[2];
    const double ul = dsqrt(ux * ux + uy * uy + uz * uz);
    for (int g = 0; g < FB; g++) {
      double pf[4];
      B.plane(g, pf);
      side[g] = pf[0] * ax + pf[1] * ay + pf[2] * az;
    }
    for (int j = 0; j < EB; j++) {
      if (!reach_b[j]) continue;
      const double cba = side[B.sides[2 * j]], dba = side[B.sides[2 * j + 1]];
      if (!(cba * dba < 0)) continue;
      double br[3];
      B.arc(j, br);
      const double bx = br[0], by = br[1], bz = br[2];
      const double adc = l[0] * bx + l[1] * by + l[2] * bz, bdc = r[0] * bx + r[1] * by + r[2] * bz;
      if (!(adc * bdc < 0 && cba * bdc < 0)) continue;
      double p[3], q[3];
      B.vert(B.edges[2 * j], p);
      B.vert(B.edges[2 * j + 1], q);
      const double tx = q[0] - p[0], ty = q[1] - p[1], tz = q[2] - p[2];
      const double len = normalize(S, uy * tz - uz * ty, uz * tx - ux * tz, ux * ty - uy * tx);
      if (!(len > PARALLEL * ul * dsqrt(tx * tx + ty * ty + tz * tz))) continue;
      double nx = S.unit[0], ny = S.unit[1], nz = S.unit[2];
      if (nx * (u[0] - ca[0]) + ny * (u[1] - ca[1]) + nz * (u[2] - ca[2]) < 0) { nx = -nx; ny = -ny; nz = -nz; }
      const double sep = nx * (p[0] - u[0]) + ny * (p[1] - u[1]) + nz * (p[2] - u[2]);
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
__device__ __forceinline__ void sphere_hull(Scratch& S, const Shape& a, const Shape& b, double margin, Manifold& out) {
  Placed B;
  B.set(b);
  const double* c = a.center;
  const double r = a.half[0];
  double most = -__builtin_inf();
  int f = 0;
  for (int k = 0; k < B.F; k++) {
    double pf[4];
    B.plane(k, pf);
    const double sep = pf[0] * c[0] + pf[1] * c[1] + pf[2] * c[2] - pf[3];
    if (sep > most) { most = sep; f = k; }
  }
  if (most - r > margin) return;
  if (most > 0 && nearest(S, a, b, r, margin, out) > 0) return;
  double pf[4];
  B.plane(f, pf);
  const double nx = -pf[0], ny = -pf[1], nz = -pf[2];
  const double separation = most - r, t = r + separation * 0.5;
  set_normal(out, nx, ny, nz);
  add_point(out, c[0] + nx * t, c[1] + ny * t, c[2] + nz * t, separation, mix(4, f, 0, 0));
}

// src/collide.js:1855-1882 (clipToHullFace), the capsule's core in S.seg_a.
__device__ __forceinline__ bool clip_to_hull_face(Scratch& S, const Placed& B, int f, double r, double margin,
                                                  Manifold& out) {
  const double* sa = S.seg_a;
  double pf[4];
  B.plane(f, pf);
  const double nx = pf[0], ny = pf[1], nz = pf[2];
  const int i0 = B.start[f], k = B.start[f + 1] - i0;
  double u0 = 0, u1 = 1;
  for (int j = 0; j < k; j++) {
    double sv[3], q[3];
    B.vert(B.loops[i0 + j], sv);
    B.vert(B.loops[i0 + (j + 1 < k ? j + 1 : 0)], q);
    const double ex = q[0] - sv[0], ey = q[1] - sv[1], ez = q[2] - sv[2];
    const double mx = ey * nz - ez * ny, my = ez * nx - ex * nz, mz = ex * ny - ey * nx;
    const double offset = mx * sv[0] + my * sv[1] + mz * sv[2];
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
    const double above = x * nx + y * ny + z * nz - pf[3], separation = above - r;
    if (separation > margin) continue;
    const double t = (above + r) * 0.5;
    add_point(out, x - nx * t, y - ny * t, z - nz * t, separation, mix(6, f, e, 0));
  }
  return out.count > 0;
}

// src/collide.js:1790-1847 (capsuleHull)
__device__ __forceinline__ void capsule_hull(Scratch& S, const Shape& a, const Shape& b, double margin,
                                             double tolerance, Manifold& out) {
  Placed B;
  B.set(b);
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
    double most = -__builtin_inf();
    for (int k = 0; k < B.F; k++) {
      double pf[4];
      B.plane(k, pf);
      const double dot = -(pf[0] * nx + pf[1] * ny + pf[2] * nz);
      if (dot > most) { most = dot; f = k; }
    }
    if (1 - most <= PARALLEL && clip_to_hull_face(S, B, f, r, margin, out)) return;
    nearest(S, a, b, r, margin, out);
    return;
  }
  const double dx = sa[3] - sa[0], dy = sa[4] - sa[1], dz = sa[5] - sa[2];
  const double dl = dsqrt(dx * dx + dy * dy + dz * dz);
  double best_sep = -__builtin_inf(), nx = 0, ny = 0, nz = 0;
  int best_face = -1, edge_i = -1;
  for (int k = 0; k < B.F; k++) {
    double pf[4];
    B.plane(k, pf);
    const double mx = pf[0], my = pf[1], mz = pf[2];
    const double sep = js_min(sa[0] * mx + sa[1] * my + sa[2] * mz, sa[3] * mx + sa[4] * my + sa[5] * mz) - pf[3] - r;
    if (sep > best_sep) { best_sep = sep; best_face = k; }
  }
  for (int i = 0; i < B.E; i++) {
    double u[3], w[3];
    B.vert(B.edges[2 * i], u);
    B.vert(B.edges[2 * i + 1], w);
    const double tx = w[0] - u[0], ty = w[1] - u[1], tz = w[2] - u[2];
    const double l = normalize(S, dy * tz - dz * ty, dz * tx - dx * tz, dx * ty - dy * tx);
    if (!(l > PARALLEL * dl * dsqrt(tx * tx + ty * ty + tz * tz))) continue;
    double mx = S.unit[0], my = S.unit[1], mz = S.unit[2];
    if (mx * (u[0] - B.c[0]) + my * (u[1] - B.c[1]) + mz * (u[2] - B.c[2]) < 0) { mx = -mx; my = -my; mz = -mz; }
    const double top = -B.lowest(-mx, -my, -mz);
    if (top - (mx * u[0] + my * u[1] + mz * u[2]) > tolerance) continue;
    const double sep = js_min(sa[0] * mx + sa[1] * my + sa[2] * mz, sa[3] * mx + sa[4] * my + sa[5] * mz) - r - top;
    if (sep > best_sep + tolerance) { best_sep = sep; edge_i = i; nx = mx; ny = my; nz = mz; }
  }
  if (best_sep > margin) return;
  if (edge_i < 0) {
    clip_to_hull_face(S, B, best_face, r, margin, out);
    return;
  }
  B.vert(B.edges[2 * edge_i], sb);
  B.vert(B.edges[2 * edge_i + 1], sb + 3);
  closest_segments(S, sa, sb);
  const double px = sa[0] + dx * S.seg_s, py = sa[1] + dy * S.seg_s, pz = sa[2] + dz * S.seg_s;
  const double qx = sb[0] + (sb[3] - sb[0]) * S.seg_t, qy = sb[1] + (sb[4] - sb[1]) * S.seg_t, qz = sb[2] + (sb[5] - sb[2]) * S.seg_t;
  const double separation = (px - qx) * nx + (py - qy) * ny + (pz - qz) * nz - r;
  set_normal(out, -nx, -ny, -nz);
  const double t = r + separation * 0.5;
  add_point(out, px - nx * t, py - ny * t, pz - nz * t, separation, mix(5, edge_i, 0, 0));
}

// src/collide.js:204-242 (collide, pair), for spheres and boxes.
// HULLS: the pairs with a hull (or a box against a hull); the rest, without. Each kernel is
// built for one kind of pair, so the other kind's working arrays take no per-lane memory.
template <bool HULLS>
__device__ __forceinline__ int collide(const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  Scratch S;
  out.count = 0;
  out.loss = 0;
  S.shape_a = &a;
  S.shape_b = &b;
  const bool swap = a.kind > b.kind;
  const Shape& lo = swap ? b : a;
  const Shape& hi = swap ? a : b;
  if constexpr (!HULLS) {
    if (lo.kind == SPHERE && hi.kind == SPHERE) sphere_sphere(S, lo, hi, margin, out);
    else if (lo.kind == SPHERE && hi.kind == CAPSULE) sphere_capsule(S, lo, hi, margin, out);
    else if (lo.kind == SPHERE && hi.kind == BOX) sphere_box(S, lo, hi, margin, out);
    else if (lo.kind == CAPSULE && hi.kind == CAPSULE) capsule_capsule(S, lo, hi, margin, tolerance, out);
    else if (lo.kind == CAPSULE && hi.kind == BOX) capsule_box(S, lo, hi, margin, tolerance, out);
    else if (lo.kind == BOX && hi.kind == BOX) box_box(S, lo, hi, margin, tolerance, out);
    else return -1;
  } else {
    if (lo.kind == SPHERE && hi.kind == HULL) sphere_hull(S, lo, hi, margin, out);
    else if (lo.kind == CAPSULE && hi.kind == HULL) capsule_hull(S, lo, hi, margin, tolerance, out);
    else if ((lo.kind == BOX || lo.kind == HULL) && hi.kind == HULL) hull_hull(S, lo, hi, margin, tolerance, out);
    else return -1;
  }
  if (swap) { out.normal[0] = -out.normal[0]; out.normal[1] = -out.normal[1]; out.normal[2] = -out.normal[2]; }
  return out.count;
}


// ---- the distance between cores, src/gjk.js ----

constexpr double PROGRESS = 1.0 / 1073741824;   // 2^-30, src/gjk.js:47
constexpr int GJK_ROUNDS = 64;                   // src/gjk.js:50
constexpr double FLAT = 1.0 / 4503599627370496;  // 2^-52, src/gjk.js:57

// src/gjk.js:60-61 (FACES): each face's corners i, j, k and the fourth corner, a nibble each
// (entry 4f + c at bits 4(4f + c)), so the table stays in registers.
__device__ __forceinline__ int face_corner(int f, int c) {
  return (int)((0x0321132023103210ull >> (16 * f + 4 * c)) & 0xf);
}

__device__ __forceinline__ double radius(const Shape& s) { return s.kind == SPHERE || s.kind == CAPSULE ? s.half[0] : 0; }

// src/gjk.js:70-107 (support) for a sphere's, a capsule's or a box's core.
__device__ __forceinline__ void gjk_support(const Shape& s, double x, double y, double z, double* out) {
  const double* c = s.center;
  const double* a = s.axes;
  const double* h = s.half;
  if (s.kind == SPHERE) {
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
  } else if (s.kind == 4) {
    // A triangle (src/collide.js:49 TRIANGLE): the corner farthest along; the first of equals.
    const double* p = s.points;
    int best = 0;
    double most = -__builtin_inf();
    for (int i = 0; i < 3; i++) {
      const double d = p[3 * i] * x + p[3 * i + 1] * y + p[3 * i + 2] * z;
      if (d > most) { most = d; best = i; }
    }
    out[0] = p[3 * best]; out[1] = p[3 * best + 1]; out[2] = p[3 * best + 2];
  } else if (s.kind == HULL) {
    const int* info = s.table->info + 8 * s.hull;
    const double* v = s.table->vertices + 3 * info[0];
    const int V = info[5];
    const double lx = x * a[0] + y * a[1] + z * a[2], ly = x * a[3] + y * a[4] + z * a[5], lz = x * a[6] + y * a[7] + z * a[8];
    int best = 0;
    double most = -__builtin_inf();
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

  __device__ __forceinline__ void move(int to, int from) {
    if (to == from) return;
    for (int c = 0; c < 3; c++) {
      W[3 * to + c] = W[3 * from + c]; PA[3 * to + c] = PA[3 * from + c]; PB[3 * to + c] = PB[3 * from + c];
    }
  }
  __device__ __forceinline__ void keep(int i, int j, int k, double li, double lj, double lk, int n) {
    move(0, i); L[0] = li;
    if (n > 1) { move(1, j); L[1] = lj; }
    if (n > 2) { move(2, k); L[2] = lk; }
    count = n;
  }
  __device__ __forceinline__ void segment(int i, int j) {
    const double ax = W[3 * i], ay = W[3 * i + 1], az = W[3 * i + 2];
    const double ex = W[3 * j] - ax, ey = W[3 * j + 1] - ay, ez = W[3 * j + 2] - az;
    const double ee = ex * ex + ey * ey + ez * ez;
    const double t = ee > 0 ? -(ax * ex + ay * ey + az * ez) / ee : 0;
    if (t <= 0) keep(i, i, i, 1, 0, 0, 1);
    else if (t >= 1) keep(j, j, j, 1, 0, 0, 1);
    else keep(i, j, j, 1 - t, t, 0, 2);
  }
  // The closest point of a triangle to the origin, by its Voronoi regions (Ericson 5.1.5).
  __device__ __forceinline__ void triangle(int i, int j, int k) {
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
  // The tetrahedron's face nearest the origin, or false when the origin is inside it.
  __device__ __forceinline__ bool tetrahedron() {
    double best = __builtin_inf();
    int best_face = -1;
    double w[12], pa[12], pb[12];
    for (int i = 0; i < 12; i++) { w[i] = W[i]; pa[i] = PA[i]; pb[i] = PB[i]; }
    bool outside = false;
    for (int f = 0; f < 4; f++) {
      const int i = face_corner(f, 0), j = face_corner(f, 1), k = face_corner(f, 2), o = face_corner(f, 3);
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
      for (int m = 0; m < 12; m++) { W[m] = w[m]; PA[m] = pa[m]; PB[m] = pb[m]; }
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
    const int i = face_corner(best_face, 0), j = face_corner(best_face, 1), k = face_corner(best_face, 2);
    for (int m = 0; m < 12; m++) { W[m] = w[m]; PA[m] = pa[m]; PB[m] = pb[m]; }
    count = 4;
    triangle(i, j, k);
    return true;
  }
  __device__ __forceinline__ bool reduce() {
    if (count == 2) { segment(0, 1); return true; }
    if (count == 3) { triangle(0, 1, 2); return true; }
    return tetrahedron();
  }
};

// src/gjk.js:114-155 (closest): the distance between two cores and the nearest points on each.
__device__ __forceinline__ double gjk_closest(const Shape& a, const Shape& b, double* near_a, double* near_b) {
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
  return dsqrt(vx * vx + vy * vy + vz * vz);
}

// ---- the time of impact, src/impact.js ----

constexpr int IMPACT_ROUNDS = 32;                  // src/impact.js:31
constexpr double UNSURE = -2;                      // src/impact.js:34
constexpr double SPIN_CAP = 3.141592653589793;     // pi, src/impact.js:28

// What meet reads of a body besides its shape: pose, motion and gravity.
struct Motion {
  double q[4], v[3], w[3];
  double gravity_scale;  // 0 unless dynamic (src/contact.js:687)
  double reach;
  bool capped;           // dynamic and not a plain sphere: spin cut to pi/dt (src/impact.js:99)
};

// src/impact.js:91-111 (start), both bodies: positions, velocities, gravity, spins, rotations.
struct Flight {
  double P[6], V[6], G[6], WV[6], Q[8];
};

__device__ __forceinline__ void start(Flight& F, const Shape& s, const Motion& m, int o, double dt) {
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
      const double k = most / dsqrt(w2);
      F.WV[o] *= k; F.WV[o + 1] *= k; F.WV[o + 2] *= k;
    }
  }
  const int q4 = o == 0 ? 0 : 4;
  for (int k = 0; k < 4; k++) F.Q[q4 + k] = m.q[k];
}

// src/impact.js:113-149 (posed, turned): where body o is at time t, falling and turning.
__device__ __forceinline__ void posed(const Flight& F, int o, double t, double* c, double* q) {
  c[0] = F.P[o] + (F.V[o] + 0.5 * F.G[o] * t) * t;
  c[1] = F.P[o + 1] + (F.V[o + 1] + 0.5 * F.G[o + 1] * t) * t;
  c[2] = F.P[o + 2] + (F.V[o + 2] + 0.5 * F.G[o + 2] * t) * t;
  const int q4 = o == 0 ? 0 : 4;
  const double wx = F.WV[o], wy = F.WV[o + 1], wz = F.WV[o + 2], w = dsqrt(wx * wx + wy * wy + wz * wz);
  if (w == 0 || t == 0) {
    for (int k = 0; k < 4; k++) q[k] = F.Q[q4 + k];
    return;
  }
  const double half = 0.5 * w * t, k = js_sin(half) / w;
  const double x = wx * k, y = wy * k, z = wz * k, cw = js_cos(half);
  const double bx = F.Q[q4], by = F.Q[q4 + 1], bz = F.Q[q4 + 2], bw = F.Q[q4 + 3];
  const double rx = cw * bx + x * bw + y * bz - z * by;
  const double ry = cw * by + y * bw + z * bx - x * bz;
  const double rz = cw * bz + z * bw + x * by - y * bx;
  const double rw = cw * bw - x * bx - y * by - z * bz;
  const double l = 1 / dsqrt(rx * rx + ry * ry + rz * rz + rw * rw);
  q[0] = rx * l; q[1] = ry * l; q[2] = rz * l; q[3] = rw * l;
}

// src/collide.js:82-91 (Shape.place)
__device__ __forceinline__ void place(Shape& s, const double* c, const double* q) {
  s.center[0] = c[0]; s.center[1] = c[1]; s.center[2] = c[2];
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  double* a = s.axes;
  a[0] = 1 - 2 * (y * y + z * z); a[1] = 2 * (x * y + z * w); a[2] = 2 * (x * z - y * w);
  a[3] = 2 * (x * y - z * w); a[4] = 1 - 2 * (x * x + z * z); a[5] = 2 * (y * z + x * w);
  a[6] = 2 * (x * z + y * w); a[7] = 2 * (y * z - x * w); a[8] = 1 - 2 * (x * x + y * y);
}

// src/impact.js:63-89 (impact): conservative advancement of the cores over [0, dt); sa and
// sb are left posed at the last time tried.
__device__ __forceinline__ double impact(Flight& F, Shape& sa, Shape& sb, const Motion& ma, const Motion& mb, double dt,
                                         double tolerance) {
  start(F, sa, ma, 0, dt);
  start(F, sb, mb, 3, dt);
  const double ra = radius(sa), rb = radius(sb);
  const double spin = dsqrt(F.WV[0] * F.WV[0] + F.WV[1] * F.WV[1] + F.WV[2] * F.WV[2]) * ma.reach
      + dsqrt(F.WV[3] * F.WV[3] + F.WV[4] * F.WV[4] + F.WV[5] * F.WV[5]) * mb.reach;
  const double gx = F.G[3] - F.G[0], gy = F.G[4] - F.G[1], gz = F.G[5] - F.G[2];
  const double gl = dsqrt(gx * gx + gy * gy + gz * gz);
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
__device__ __forceinline__ double moving(const Motion& ma, const Motion& mb, const double* g, double dt) {
  const double dx = mb.v[0] - ma.v[0], dy = mb.v[1] - ma.v[1], dz = mb.v[2] - ma.v[2];
  const double gs = mb.gravity_scale - ma.gravity_scale;
  const double gl = dsqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) * dabs(gs);
  const double turning = dsqrt(ma.w[0] * ma.w[0] + ma.w[1] * ma.w[1] + ma.w[2] * ma.w[2]) * ma.reach
      + dsqrt(mb.w[0] * mb.w[0] + mb.w[1] * mb.w[1] + mb.w[2] * mb.w[2]) * mb.reach;
  return (dsqrt(dx * dx + dy * dy + dz * dz) + 0.5 * gl * dt + turning) * dt;
}

// src/contact.js:740-761 (carried), a whole body: a point held body-relative at the moment of
// impact (turned qt) carried to the step's start (turned q).
__device__ __forceinline__ V3 carried(const Shape& s, const double* qt, const double* q, double x, double y, double z) {
  if (s.kind == SPHERE) return V3{x, y, z};
  V3 o = rotate(qt, x, y, z, true);
  if (s.kind != CAPSULE) return rotate(q, o.x, o.y, o.z, false);
  const double half = s.half[1];
  const double c = o.y < -half ? -half : o.y > half ? half : o.y;
  const V3 u = rotate(qt, 0, c, 0, false);
  const double rx = x - u.x, ry = y - u.y, rz = z - u.z;
  V3 r = rotate(q, 0, c, 0, false);
  r.x += rx; r.y += ry; r.z += rz;
  return r;
}

// src/contact.js:712-732 (back): the manifold found at the moment of impact, carried back to
// the step's start with b's end's offset from a's kept in shift.
__device__ __forceinline__ void back(const Flight& F, const Shape& sa, const Shape& sb, const Motion& ma,
                                     const Motion& mb, double t, Manifold& mf, double* shift) {
  double CA[3], QA[4], CB[3], QB[4];
  posed(F, 0, t, CA, QA);
  posed(F, 3, t, CB, QB);
  const double nx = mf.normal[0], ny = mf.normal[1], nz = mf.normal[2];
  for (int k = 0; k < mf.count; k++) {
    const int k3 = 3 * k;
    const double half = 0.5 * mf.separations[k];
    const double px = mf.points[k3], py = mf.points[k3 + 1], pz = mf.points[k3 + 2];
    const V3 ea = carried(sa, QA, ma.q, px - nx * half - CA[0], py - ny * half - CA[1], pz - nz * half - CA[2]);
    const double ax = F.P[0] + ea.x, ay = F.P[1] + ea.y, az = F.P[2] + ea.z;
    const V3 eb = carried(sb, QB, mb.q, px + nx * half - CB[0], py + ny * half - CB[1], pz + nz * half - CB[2]);
    const double bx = F.P[3] + eb.x - ax, by = F.P[4] + eb.y - ay, bz = F.P[5] + eb.z - az;
    mf.separations[k] = bx * nx + by * ny + bz * nz;
    mf.points[k3] = ax; mf.points[k3 + 1] = ay; mf.points[k3 + 2] = az;
    shift[k3] = bx; shift[k3 + 1] = by; shift[k3 + 2] = bz;
  }
}

__device__ __forceinline__ double least_separation(const Manifold& m) {
  double s = __builtin_inf();
  for (int k = 0; k < m.count; k++)
    if (m.separations[k] < s) s = m.separations[k];
  return s;
}

// src/contact.js:627-640 (meet) for two whole bodies. Returns the point count (0: none),
// or -1 for a pair of kinds not handled; flags 1 shifted, 2 a time of impact asked.
template <bool HULLS>
__device__ __forceinline__ int meet(Shape& sa, Shape& sb, const Motion& ma, const Motion& mb, const double* gravity,
                                    double dt, double margin, double tolerance, Manifold& out, double* shift,
                                    int& flags) {
  flags = 0;
  // Two rounds of one collide() call: the step's start, then (after a time of impact) the
  // moment of impact with the shapes placed there, so the dispatch and its scratch are
  // inlined once.
  Flight F;
  double t = 0, grow = margin;
  for (int round = 0;; round++) {
    const int n = collide<HULLS>(sa, sb, grow, tolerance, out);
    if (n <= 0) return n;
    if (round == 1) {
      back(F, sa, sb, ma, mb, t, out, shift);
      flags |= 1;
      return out.count;
    }
    const double moves = least_separation(out) > tolerance ? moving(ma, mb, gravity, dt) : 0;
    if (!(moves > tolerance)) return out.count;
    t = impact(F, sa, sb, ma, mb, dt, tolerance);
    flags |= 2;
    grow = t >= 0 ? moves * (1 - t / dt) + 2 * tolerance : 0;
    if (t == -1) return 0;
    if (t == UNSURE) return out.count;
  }
}


#include "level_dev.h"

}  // namespace narrow

// One pair a lane: shape records (kind, centre 3, axes 9, half 3, weight 3) in, manifolds out;
// the pairs with a hull in the HULLS kernel, the others in the other one.
__device__ __forceinline__ bool has_hull(const double* a, const double* b) {
  return (int)a[0] == narrow::HULL || (int)b[0] == narrow::HULL;
}

template <bool HULLS>
__global__ void collide_kernel(CollideDev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.n || has_hull(d.a + 19 * (long)k, d.b + 19 * (long)k) != HULLS) return;
  narrow::Shape s[2];
  const double* src[2] = {d.a + 19 * (long)k, d.b + 19 * (long)k};
  for (int t = 0; t < 2; t++) {
    const double* r = src[t];
    s[t].kind = (int)r[0];
    for (int i = 0; i < 3; i++) s[t].center[i] = r[1 + i];
    for (int i = 0; i < 9; i++) s[t].axes[i] = r[4 + i];
    for (int i = 0; i < 3; i++) s[t].half[i] = r[13 + i];
    for (int i = 0; i < 3; i++) s[t].weight[i] = r[16 + i];
    s[t].hull = s[t].kind == narrow::HULL ? (int)r[13] : -1;
    s[t].table = &d.hulls;
  }
  narrow::Manifold m;
  m.count = 0;
  const int n = narrow::collide<HULLS>(s[0], s[1], d.margin[k], d.tolerance[k], m);
  d.count[k] = n;
  for (int i = 0; i < 3; i++) d.normal[3 * k + i] = m.normal[i];
  for (int i = 0; i < 3 * n; i++) d.points[12 * (long)k + i] = m.points[i];
  for (int i = 0; i < n; i++) d.separations[4 * (long)k + i] = m.separations[i];
  for (int i = 0; i < n; i++) d.ids[4 * (long)k + i] = m.ids[i];
  d.loss[k] = m.loss;
}

// One pair a lane: body records (32 numbers) and pair records (margin, tolerance, dt,
// gravity 3) in; meet's manifold, shift and flags out.
template <bool HULLS>
__global__ void meet_kernel(MeetDev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.n) return;
  const double* src[2] = {d.a + 32 * (long)(d.ia ? d.ia[k] : k), d.b + 32 * (long)(d.ib ? d.ib[k] : k)};
  if (has_hull(src[0], src[1]) != HULLS) return;
  narrow::Shape s[2];
  narrow::Motion mo[2];
  for (int t = 0; t < 2; t++) {
    const double* r = src[t];
    s[t].kind = (int)r[0];
    for (int i = 0; i < 3; i++) s[t].center[i] = r[1 + i];
    for (int i = 0; i < 9; i++) s[t].axes[i] = r[4 + i];
    for (int i = 0; i < 3; i++) s[t].half[i] = r[13 + i];
    for (int i = 0; i < 3; i++) s[t].weight[i] = r[16 + i];
    for (int i = 0; i < 4; i++) mo[t].q[i] = r[19 + i];
    for (int i = 0; i < 3; i++) mo[t].v[i] = r[23 + i];
    for (int i = 0; i < 3; i++) mo[t].w[i] = r[26 + i];
    s[t].hull = s[t].kind == narrow::HULL ? (int)r[13] : -1;
    s[t].table = &d.hulls;
    mo[t].gravity_scale = r[29];
    mo[t].reach = r[30];
    mo[t].capped = (((int)r[31]) & 2) != 0;
  }
  const double* pr = d.pair + 6 * (long)k;
  narrow::Manifold m;
  m.count = 0;
  for (int i = 0; i < 3; i++) m.normal[i] = 0;
  double shift[12];
  for (int i = 0; i < 12; i++) shift[i] = 0;
  int flags = 0;
  const int n = narrow::meet<HULLS>(s[0], s[1], mo[0], mo[1], pr + 3, pr[2], pr[0], pr[1], m, shift, flags);
  d.count[k] = n;
  d.flags[k] = flags;
  for (int i = 0; i < 3; i++) d.normal[3 * k + i] = m.normal[i];
  for (int i = 0; i < 3 * n; i++) d.points[12 * (long)k + i] = m.points[i];
  for (int i = 0; i < n; i++) d.separations[4 * (long)k + i] = m.separations[i];
  for (int i = 0; i < n; i++) d.ids[4 * (long)k + i] = m.ids[i];
  for (int i = 0; i < 12; i++) d.shift[12 * (long)k + i] = shift[i];
  d.loss[k] = m.loss;
}

// touchLevel, one shape a lane (src/triangle.js:1078-1131, visit 1018-1069): the triangles
// of every cell under the box whose heights reach it, row by row, each touched; points
// grouped by direction, kept once where two triangles give the same one; each group sorted
// by id and reduced to four.
__global__ void __launch_bounds__(64, 2) level_kernel(LevelDev d) {
  using namespace narrow;
  __shared__ Touch touched[64];
  __shared__ int hits[64];
  const int k = blockIdx.x, lane = threadIdx.x;
  if (k >= d.n) return;
  const double* r = d.shapes + 19 * (long)k;
  Shape s;
  s.kind = (int)r[0];
  for (int i = 0; i < 3; i++) s.center[i] = r[1 + i];
  for (int i = 0; i < 9; i++) s.axes[i] = r[4 + i];
  for (int i = 0; i < 3; i++) s.half[i] = r[13 + i];
  for (int i = 0; i < 3; i++) s.weight[i] = r[16 + i];
  s.hull = s.kind == HULL ? (int)r[13] : -1;
  s.table = &d.hulls;
  s.points = nullptr;
  const int f = d.field[k];
  Field F;
  F.columns = d.fields.info[2 * f];
  F.rows = d.fields.info[2 * f + 1];
  F.h = d.fields.heights + d.fields.off[f];
  F.spacing = d.fields.dinfo[3 * f];
  F.x0 = d.fields.dinfo[3 * f + 1];
  F.z0 = d.fields.dinfo[3 * f + 2];
  const double* c = d.position + 3 * (long)k;
  const double* A = d.axes + 9 * (long)k;
  const double* box = d.boxes + 6 * (long)k;
  const double margin = d.margin[k], tolerance = d.tolerance[k];
  const long long base = d.scratch[k];
  const int cap_points = (int)(d.scratch[k + 1] - base);
  LevelScratch W;
  W.at = d.sd + 9 * base;
  W.gat = W.at + 4 * (long)cap_points;
  W.normals = W.gat + 4 * (long)cap_points;
  W.id = d.su + 3 * base;
  W.in = W.id + cap_points;
  W.gid = W.in + cap_points;
  Toucher T;
  Touch mine;
  int found = 0, groups = 0;
  const double t2 = tolerance * tolerance;
  // src/mesh.js:388-401 (triangles)
  const int cells = F.columns - 1;
  const double i0 = narrow::js_max(0, floor((box[0] - F.x0) / F.spacing)), i1 = narrow::js_min(cells - 1, floor((box[3] - F.x0) / F.spacing));
  const double j0 = narrow::js_max(0, floor((box[2] - F.z0) / F.spacing)), j1 = narrow::js_min(F.rows - 2, floor((box[5] - F.z0) / F.spacing));
  if (i0 <= i1 && j0 <= j1) {
    // A wave a pair: lane l touches triangle start + l of the serial walk's order (rows of
    // cells, cells, then the cell's two halves), each touch its own (touch() starts afresh),
    // and the first lane takes the 64 touches in that order into the groups, as the walk does.
    const int width = (int)i1 - (int)i0 + 1, total = 2 * width * ((int)j1 - (int)j0 + 1);
    for (int start = 0; start < total; start += 64) {
      const int o = start + lane;
      int got = 0;
      if (o < total) {
        const int cell = o >> 1, half = o & 1;
        const int i = (int)i0 + cell % width, j = (int)j0 + cell / width;
        const double ha = F.at(i, j), hb = F.at(i + 1, j), hc = F.at(i, j + 1), hd = F.at(i + 1, j + 1);
        if (!(narrow::js_max(narrow::js_max(narrow::js_max(ha, hb), hc), hd) < box[1] || narrow::js_min(narrow::js_min(narrow::js_min(ha, hb), hc), hd) > box[4])) {
          const int t = 2 * (j * cells + i) + half;
          double local[9], p[9];
          F.triangle(t, local);
          for (int v = 0; v < 3; v++) {
            const double x = local[3 * v], y = local[3 * v + 1], z = local[3 * v + 2];
            p[3 * v] = c[0] + A[0] * x + A[3] * y + A[6] * z;
            p[3 * v + 1] = c[1] + A[1] * x + A[4] * y + A[7] * z;
            p[3 * v + 2] = c[2] + A[2] * x + A[5] * y + A[8] * z;
          }
          Shape tri;
          tri.kind = TRIANGLE;
          tri.points = p;
          tri.hull = -1;
          tri.table = nullptr;
          for (int q = 0; q < 9; q++) tri.axes[q] = (q == 0 || q == 4 || q == 8) ? 1 : 0;
          for (int q = 0; q < 3; q++) { tri.half[q] = 0; tri.weight[q] = 0; }
          tri.center[0] = (p[0] + p[3] + p[6]) / 3;
          tri.center[1] = (p[1] + p[4] + p[7]) / 3;
          tri.center[2] = (p[2] + p[5] + p[8]) / 3;
          triangle_normal(p, tri.axes);
          const int flags = field_edge(F, t, T.folds);
          got = T.touch(s, tri, flags, margin, tolerance, mine, t);
        }
      }
      hits[lane] = got;
      if (got) touched[lane] = mine;
      __syncthreads();
      if (lane == 0) {
        const int batch = total - start < 64 ? total - start : 64;
        for (int l = 0; l < batch; l++) {
          if (hits[l] == 0) continue;
          const Touch& one = touched[l];
          const double nx = one.normal[0], ny = one.normal[1], nz = one.normal[2];
          int g = 0;
          while (g < groups && W.normals[3 * g] * nx + W.normals[3 * g + 1] * ny + W.normals[3 * g + 2] * nz < 1 - SAME) g++;
          if (g == groups) {
            W.normals[3 * g] = nx; W.normals[3 * g + 1] = ny; W.normals[3 * g + 2] = nz;
            groups++;
          }
          for (int q = 0; q < one.count; q++) {
            const double x = one.points[3 * q], y = one.points[3 * q + 1], z = one.points[3 * q + 2], sep = one.separations[q];
            int same = -1;
            for (int e = 0; e < found && same < 0; e++) {
              if ((int)W.in[e] != g) continue;
              const double dx = W.at[4 * e] - x, dy = W.at[4 * e + 1] - y, dz = W.at[4 * e + 2] - z;
              if (dx * dx + dy * dy + dz * dz <= t2) same = e;
            }
            if (same >= 0) {
              if (sep < W.at[4 * same + 3]) {
                W.at[4 * same] = x; W.at[4 * same + 1] = y; W.at[4 * same + 2] = z; W.at[4 * same + 3] = sep;
              }
              if (one.ids[q] < 512 && W.id[same] >= 512) W.id[same] = one.ids[q];
              continue;
            }
            if (found == cap_points) continue;   // never: at most 7 points a triangle
            W.at[4 * found] = x; W.at[4 * found + 1] = y; W.at[4 * found + 2] = z; W.at[4 * found + 3] = sep;
            W.id[found] = one.ids[q];
            W.in[found] = (uint32_t)g;
            found++;
          }
        }
      }
      __syncthreads();
    }
  }
  if (lane != 0) return;
  for (int g = 0; g < groups; g++) {
    int m = 0;
    for (int i = 0; i < found; i++) {
      if ((int)W.in[i] != g) continue;
      int j = m++;
      for (; j > 0 && W.gid[j - 1] > W.id[i]; j--) {
        for (int q = 0; q < 4; q++) W.gat[4 * j + q] = W.gat[4 * j - 4 + q];
        W.gid[j] = W.gid[j - 1];
      }
      for (int q = 0; q < 4; q++) W.gat[4 * j + q] = W.at[4 * i + q];
      W.gid[j] = W.id[i];
    }
    Manifold mf;
    mf.count = 0;
    mf.normal[0] = W.normals[3 * g]; mf.normal[1] = W.normals[3 * g + 1]; mf.normal[2] = W.normals[3 * g + 2];
    level_reduce(W, m, tolerance, mf);
    if (g >= d.cap) continue;
    const long o = (long)k * d.cap + g;
    d.count[o] = mf.count;
    for (int q = 0; q < 3; q++) d.normal[3 * o + q] = mf.normal[q];
    for (int q = 0; q < 12; q++) d.points[12 * o + q] = q < 3 * mf.count ? mf.points[q] : 0;
    for (int q = 0; q < 4; q++) d.separations[4 * o + q] = q < mf.count ? mf.separations[q] : 0;
    for (int q = 0; q < 4; q++) d.ids[4 * o + q] = q < mf.count ? mf.ids[q] : 0;
  }
  d.manifolds[k] = groups;
}

void launch_level(const LevelDev& d, hipStream_t stream) {
  if (d.n == 0) return;
  hipLaunchKernelGGL(level_kernel, dim3(d.n), dim3(64), 0, stream, d);
}

// The pass's results packed by pair (LevelPacked): first every pair's share, min(manifolds, cap),
// scanned by one workgroup (each thread a run of pairs, then the runs' sums in LDS), then a lane
// a pair copying its slots to their packed rows.
__global__ void __launch_bounds__(1024) level_scan_kernel(LevelDev d, int* start) {
  __shared__ int sums[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (d.n + T - 1) / T, k0 = t * per, k1 = k0 + per < d.n ? k0 + per : d.n;
  int run = 0;
  for (int k = k0; k < k1; k++) run += d.manifolds[k] < d.cap ? d.manifolds[k] : d.cap;
  sums[t] = run;
  __syncthreads();
  for (int step = 1; step < T; step *= 2) {
    const int add = t >= step ? sums[t - step] : 0;
    __syncthreads();
    sums[t] += add;
    __syncthreads();
  }
  int at = sums[t] - run;
  for (int k = k0; k < k1; k++) {
    start[k] = at;
    at += d.manifolds[k] < d.cap ? d.manifolds[k] : d.cap;
  }
  if (t == T - 1) start[d.n] = sums[t];
}

__global__ void __launch_bounds__(256) level_pack_kernel(LevelDev d, LevelPacked p) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.n) return;
  const int got = d.manifolds[k] < d.cap ? d.manifolds[k] : d.cap;
  for (int j = 0; j < got; j++) {
    const long o = (long)k * d.cap + j, r = p.start[k] + j;
    p.count[r] = d.count[o];
    for (int i = 0; i < 3; i++) p.normal[3 * r + i] = d.normal[3 * o + i];
    for (int i = 0; i < 12; i++) p.points[12 * r + i] = d.points[12 * o + i];
    for (int i = 0; i < 4; i++) p.separations[4 * r + i] = d.separations[4 * o + i];
    for (int i = 0; i < 4; i++) p.ids[4 * r + i] = d.ids[4 * o + i];
  }
}

void launch_level_pack(const LevelDev& d, const LevelPacked& p, hipStream_t stream) {
  if (d.n == 0) return;
  hipLaunchKernelGGL(level_scan_kernel, dim3(1), dim3(1024), 0, stream, d, p.start);
  hipLaunchKernelGGL(level_pack_kernel, dim3((d.n + 255) / 256), dim3(256), 0, stream, d, p);
}

// ---- the level pass without allocation ----

// The block: pair offsets (LEVEL_NOALLOC_PAIRS + 1 of them), the status (3 words), then the
// found points' doubles and words.
constexpr size_t NOALLOC_OFFSETS = 8 * ((size_t)LEVEL_NOALLOC_PAIRS + 1);
constexpr size_t NOALLOC_STATUS = (NOALLOC_OFFSETS + 255) / 256 * 256;
constexpr size_t NOALLOC_SD = NOALLOC_STATUS + 256;
constexpr size_t NOALLOC_SU = NOALLOC_SD + 9 * 8 * (size_t)LEVEL_NOALLOC_POINTS;
constexpr size_t NOALLOC_BYTES = NOALLOC_SU + 3 * 4 * (size_t)LEVEL_NOALLOC_POINTS;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

hipError_t level_noalloc_block(int device, void** out) {
  static std::mutex lock;
  static std::vector<void*> at;
  std::lock_guard<std::mutex> held(lock);
  if (device < 0) return hipErrorInvalidDevice;
  if ((int)at.size() <= device) at.resize(device + 1, nullptr);
  if (!at[device]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess) return rc;
    at[device] = p;
  }
  *out = at[device];
  return hipSuccess;
}

const void* level_plan_status(const void* block) { return static_cast<const unsigned char*>(block) + NOALLOC_STATUS; }

void level_on_block(LevelDev& d, void* block) {
  unsigned char* b = static_cast<unsigned char*>(block);
  d.scratch = reinterpret_cast<const long long*>(b);
  d.sd = reinterpret_cast<double*>(b + NOALLOC_SD);
  d.su = reinterpret_cast<uint32_t*>(b + NOALLOC_SU);
}

// One workgroup: each thread takes a run of pairs, checks them as bm_level_host does and
// counts the points each may find (7 a triangle under its box, the cells the pass itself
// walks); the runs' totals are scanned in LDS, then each thread writes its pairs' offsets.
constexpr int PLAN_THREADS = 1024;
__global__ void __launch_bounds__(PLAN_THREADS) level_plan_kernel(LevelDev d, long long* offsets, long long* status) {
  __shared__ long long run[PLAN_THREADS];
  __shared__ int flags[PLAN_THREADS], first[PLAN_THREADS];
  const int t = threadIdx.x, n = d.n;
  const int per = (n + PLAN_THREADS - 1) / PLAN_THREADS;
  const int lo = t * per < n ? t * per : n, hi = lo + per < n ? lo + per : n;
  long long sum = 0;
  int bad = 0, at = n;
  for (int k = lo; k < hi; k++) {
    const double* r = d.shapes + 19 * (long)k;
    const double kind = r[0];
    int why = 0;
    if (!(kind == 0 || kind == 1 || kind == 2 || kind == 3)) why |= LEVEL_PLAN_KIND;
    if (kind == 3 && !(r[13] >= 0 && r[13] < d.hulls.n && r[13] == (double)(int)r[13])) why |= LEVEL_PLAN_HULL;
    const int f = d.field[k];
    if (f < 0 || f >= d.fields.n) why |= LEVEL_PLAN_FIELD;
    long long need = 0;
    if (!why) {
      const int columns = d.fields.info[2 * f], rows = d.fields.info[2 * f + 1];
      const double spacing = d.fields.dinfo[3 * f], x0 = d.fields.dinfo[3 * f + 1], z0 = d.fields.dinfo[3 * f + 2];
      const double* box = d.boxes + 6 * (long)k;
      // The pass's own walk (src/mesh.js:388-401): the cells under the box.
      const double i0 = narrow::js_max(0, floor((box[0] - x0) / spacing)),
                   i1 = narrow::js_min(columns - 2, floor((box[3] - x0) / spacing));
      const double j0 = narrow::js_max(0, floor((box[2] - z0) / spacing)),
                   j1 = narrow::js_min(rows - 2, floor((box[5] - z0) / spacing));
      if (i0 <= i1 && j0 <= j1) {
        const double points = 14 * (i1 - i0 + 1) * (j1 - j0 + 1);
        if (points > (double)LEVEL_NOALLOC_POINTS) why |= LEVEL_PLAN_POINTS;
        else need = (long long)points;
      }
    }
    if (why && at == n) at = k;
    bad |= why;
    offsets[k + 1] = need;
    sum += need;
  }
  run[t] = sum;
  flags[t] = bad;
  first[t] = at;
  __syncthreads();
  if (t == 0) {
    long long total = 0;
    int all = 0, earliest = n;
    for (int i = 0; i < PLAN_THREADS; i++) {
      const long long mine = run[i];
      run[i] = total;
      total += mine;
      all |= flags[i];
      if (first[i] < earliest) earliest = first[i];
    }
    if (total > LEVEL_NOALLOC_POINTS && !(all & LEVEL_PLAN_POINTS)) all |= LEVEL_PLAN_POINTS;
    offsets[0] = 0;
    status[0] = total;
    status[1] = all;
    status[2] = earliest;
  }
  __syncthreads();
  long long before = run[t];
  for (int k = lo; k < hi; k++) {
    before += offsets[k + 1];
    offsets[k + 1] = before;
  }
}

void launch_level_plan(const LevelDev& d, void* block, hipStream_t stream) {
  unsigned char* b = static_cast<unsigned char*>(block);
  hipLaunchKernelGGL(level_plan_kernel, dim3(1), dim3(PLAN_THREADS), 0, stream, d,
                     reinterpret_cast<long long*>(b), reinterpret_cast<long long*>(b + NOALLOC_STATUS));
}

__global__ void level_clear_kernel(LevelDev d) {
  const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (long)d.n * d.cap) return;
  const long k = slot / d.cap;
  const int j = (int)(slot - k * d.cap);
  const int got = d.manifolds[k] < d.cap ? d.manifolds[k] : d.cap;
  if (j < got) return;
  d.count[slot] = 0;
  for (int q = 0; q < 3; q++) d.normal[3 * slot + q] = 0;
  for (int q = 0; q < 12; q++) d.points[12 * slot + q] = 0;
  for (int q = 0; q < 4; q++) {
    d.separations[4 * slot + q] = 0;
    d.ids[4 * slot + q] = 0;
  }
}

void launch_level_clear(const LevelDev& d, hipStream_t stream) {
  const long slots = (long)d.n * d.cap;
  if (slots == 0) return;
  hipLaunchKernelGGL(level_clear_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, stream, d);
}

void launch_meet(const MeetDev& d, hipStream_t stream) {
  if (d.n == 0) return;
  hipLaunchKernelGGL(meet_kernel<false>, dim3((d.n + 63) / 64), dim3(64), 0, stream, d);
  hipLaunchKernelGGL(meet_kernel<true>, dim3((d.n + 63) / 64), dim3(64), 0, stream, d);
}

// Which pairs of a meet pass found a manifold, numbered in pair order, in three launches: each
// workgroup of 256 pairs counts its own, one workgroup turns the counts into each workgroup's
// first row (and the total), then each pair takes its row (the workgroup's first plus the pairs
// before it in the workgroup that found one) and copies its results there.
constexpr int MEET_PACK_THREADS = 256;

__global__ void __launch_bounds__(MEET_PACK_THREADS) meet_count_kernel(MeetDev d, int* sums) {
  const int k = blockIdx.x * MEET_PACK_THREADS + threadIdx.x;
  const int found = __syncthreads_count(k < d.n && d.count[k] > 0);
  if (threadIdx.x == 0) sums[blockIdx.x] = found;
}

__global__ void __launch_bounds__(1024) meet_first_kernel(int blocks, int* sums, int* found) {
  __shared__ int part[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (blocks + T - 1) / T, b0 = t * per, b1 = b0 + per < blocks ? b0 + per : blocks;
  int run = 0;
  for (int b = b0; b < b1; b++) run += sums[b];
  part[t] = run;
  __syncthreads();
  for (int step = 1; step < T; step *= 2) {
    const int add = t >= step ? part[t - step] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int at = part[t] - run;
  for (int b = b0; b < b1; b++) {
    const int n = sums[b];
    sums[b] = at;
    at += n;
  }
  if (t == T - 1) *found = part[t];
}

__global__ void __launch_bounds__(MEET_PACK_THREADS) meet_pack_kernel(MeetDev d, MeetPacked p, const int* first) {
  __shared__ int waves[MEET_PACK_THREADS / 64];
  const int k = blockIdx.x * MEET_PACK_THREADS + threadIdx.x;
  const bool has = k < d.n && d.count[k] > 0;
  const unsigned long long mask = __ballot(has);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) waves[wave] = __popcll(mask);
  __syncthreads();
  int before = first[blockIdx.x];
  for (int w = 0; w < wave; w++) before += waves[w];
  before += __popcll(mask & ((1ull << lane) - 1));
  if (k >= d.n) return;
  p.row[k] = has ? before : -1;
  if (!has) return;
  const long r = before;
  for (int i = 0; i < 3; i++) p.normal[3 * r + i] = d.normal[3 * (long)k + i];
  for (int i = 0; i < 12; i++) p.points[12 * r + i] = d.points[12 * (long)k + i];
  for (int i = 0; i < 4; i++) p.separations[4 * r + i] = d.separations[4 * (long)k + i];
  for (int i = 0; i < 12; i++) p.shift[12 * r + i] = d.shift[12 * (long)k + i];
  for (int i = 0; i < 4; i++) p.ids[4 * r + i] = d.ids[4 * (long)k + i];
  p.loss[r] = d.loss[k];
}

void launch_meet_pack(const MeetDev& d, const MeetPacked& p, hipStream_t stream) {
  if (d.n == 0) return;
  const int blocks = (d.n + MEET_PACK_THREADS - 1) / MEET_PACK_THREADS;
  hipLaunchKernelGGL(meet_count_kernel, dim3(blocks), dim3(MEET_PACK_THREADS), 0, stream, d, p.sums);
  hipLaunchKernelGGL(meet_first_kernel, dim3(1), dim3(1024), 0, stream, blocks, p.sums, p.found);
  hipLaunchKernelGGL(meet_pack_kernel, dim3(blocks), dim3(MEET_PACK_THREADS), 0, stream, d, p,
                     static_cast<const int*>(p.sums));
}

void launch_collide(const CollideDev& d, hipStream_t stream) {
  if (d.n == 0) return;
  hipLaunchKernelGGL(collide_kernel<false>, dim3((d.n + 63) / 64), dim3(64), 0, stream, d);
  hipLaunchKernelGGL(collide_kernel<true>, dim3((d.n + 63) / 64), dim3(64), 0, stream, d);
}

}  // namespace bm
