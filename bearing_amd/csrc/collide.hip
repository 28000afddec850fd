// The narrowphase on the MI355X for spheres and boxes, one pair a lane: src/collide.js
// collide / pair (204-242), the sphere pairs (349-403) and box against box (713-1229:
// fifteen separating axes, the reference face clipped with snap slack, feature ids named
// by both boxes, the four-point reduction with its HELD UP / UNDER restarts, edge
// contact, closest points). Every branch and every operation is the reference's, in its
// order, in doubles without contraction: the manifolds are the reference's bits.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

constexpr int SPHERE = 0, CAPSULE = 1, BOX = 2;
constexpr int MAX_POINTS = 4;
constexpr double PARALLEL = 7.450580596923828e-9;  // 2^-27, src/collide.js:686

struct Shape {
  int kind;
  double center[3];
  double axes[9];
  double half[3];
  double weight[3];
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
  const double* kept = S.kept;
  const double* sep = S.kept_sep;
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
  const double t = ((S.kept[3 * o] - c[0]) * nx + (S.kept[3 * o + 1] - c[1]) * ny + (S.kept[3 * o + 2] - c[2]) * nz) / along;
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
    const double ax = S.kept[3 * from] - px, ay = S.kept[3 * from + 1] - py, az = S.kept[3 * from + 2] - pz;
    const double bx = S.kept[3 * to] - px, by = S.kept[3 * to + 1] - py, bz = S.kept[3 * to + 2] - pz;
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
    if (S.kept_sep[k] > pressed) continue;
    const double x = S.kept[3 * k] - S.foot[0], y = S.kept[3 * k + 1] - S.foot[1], z = S.kept[3 * k + 2] - S.foot[2];
    const double along = x * nx + y * ny + z * nz;
    const double d = x * x + y * y + z * z - along * along;
    if (d < near) { near = d; start = k; }
  }
  return start;
}

// src/collide.js:1077-1104 (reducePoints)
__device__ __forceinline__ void reduce_points(Scratch& S, int n, double nx, double ny, double nz, double tolerance, Manifold& out) {
  const double* kept = S.kept;
  const double* sep = S.kept_sep;
  const uint32_t* id = S.kept_id;
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

// src/collide.js:204-242 (collide, pair), for spheres and boxes.
__device__ __forceinline__ int collide(const Shape& a, const Shape& b, double margin, double tolerance, Manifold& out) {
  Scratch S;
  out.count = 0;
  out.loss = 0;
  S.shape_a = &a;
  S.shape_b = &b;
  const bool swap = a.kind > b.kind;
  const Shape& lo = swap ? b : a;
  const Shape& hi = swap ? a : b;
  if (lo.kind == SPHERE && hi.kind == SPHERE) sphere_sphere(S, lo, hi, margin, out);
  else if (lo.kind == SPHERE && hi.kind == BOX) sphere_box(S, lo, hi, margin, out);
  else if (lo.kind == BOX && hi.kind == BOX) box_box(S, lo, hi, margin, tolerance, out);
  else return -1;
  if (swap) { out.normal[0] = -out.normal[0]; out.normal[1] = -out.normal[1]; out.normal[2] = -out.normal[2]; }
  return out.count;
}


}  // namespace narrow

// One pair a lane: shape records (kind, centre 3, axes 9, half 3, weight 3) in, manifolds out.
__global__ void collide_kernel(CollideDev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.n) return;
  narrow::Shape s[2];
  const double* src[2] = {d.a + 19 * (long)k, d.b + 19 * (long)k};
  for (int t = 0; t < 2; t++) {
    const double* r = src[t];
    s[t].kind = (int)r[0];
    for (int i = 0; i < 3; i++) s[t].center[i] = r[1 + i];
    for (int i = 0; i < 9; i++) s[t].axes[i] = r[4 + i];
    for (int i = 0; i < 3; i++) s[t].half[i] = r[13 + i];
    for (int i = 0; i < 3; i++) s[t].weight[i] = r[16 + i];
  }
  narrow::Manifold m;
  m.count = 0;
  const int n = narrow::collide(s[0], s[1], d.margin[k], d.tolerance[k], m);
  d.count[k] = n;
  for (int i = 0; i < 3; i++) d.normal[3 * k + i] = m.normal[i];
  for (int i = 0; i < 3 * n; i++) d.points[12 * (long)k + i] = m.points[i];
  for (int i = 0; i < n; i++) d.separations[4 * (long)k + i] = m.separations[i];
  for (int i = 0; i < n; i++) d.ids[4 * (long)k + i] = m.ids[i];
}

void launch_collide(const CollideDev& d, hipStream_t stream) {
  if (d.n == 0) return;
  hipLaunchKernelGGL(collide_kernel, dim3((d.n + 63) / 64), dim3(64), 0, stream, d);
}

}  // namespace bm
