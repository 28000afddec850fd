// Level geometry on the device: a shape against the triangles of a heightfield, one shape a
// lane. Included into collide.hip inside namespace bm::narrow, after GJK, so it shares that
// file's Shape, Placed and gjk_closest.
//
//   src/mesh.js:67-125, 288-455   flags, fold, normal; the Heightfield's cells, triangles, edges
//   src/triangle.js:104-986       touch: sphere, capsule, box and hull against one triangle
//   src/triangle.js:988-1183      touchLevel: every triangle under a box, its points grouped by
//                                 normal (within SAME), kept once when two triangles give the
//                                 same point, each group sorted by id and reduced to four
//
// The found points, the group normals and a group's gathered points live in a per-lane
// scratch block in HBM, sized by the host from the number of triangles under the lane's box
// (at most 7 points a triangle, at most one group a triangle).
#pragma once

constexpr int TRIANGLE = 4;                           // src/collide.js:49
constexpr double MESH_FLAT = 1.0 / 1048576;           // 2^-20, src/mesh.js:40
constexpr int FRONT = 1, BACK = 8, ONE_SIDED = 64;    // src/mesh.js:47, 60
constexpr int LEVEL_MOST = 7;                         // src/triangle.js:102 (MOST)
constexpr uint32_t CORNER_ID = 1u << 30;              // src/triangle.js:840 (CORNER)
constexpr double SAME = MESH_FLAT * MESH_FLAT / 2;    // src/triangle.js:986
constexpr int LEVEL_CLIP = 24;                        // clip buffer: a 16-corner face + 3 cuts, or a triangle + 16

__device__ __forceinline__ double js_max(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a > b) return a;
  if (b > a) return b;
  return __builtin_signbit(a) ? b : a;
}

// A heightfield as read on the device (src/mesh.js:288-311).
struct Field {
  const double* h;
  int columns, rows;
  double spacing, x0, z0;
  __device__ __forceinline__ double at(int i, int j) const { return h[(long)j * columns + i]; }
  __device__ __forceinline__ void grid(int t, int* g) const {
    const int cell = t >> 1, cells = columns - 1, i = cell % cells, j = (cell - i) / cells;
    if ((t & 1) == 0) { g[0] = i; g[1] = j; g[2] = i + 1; g[3] = j + 1; g[4] = i + 1; g[5] = j; }
    else { g[0] = i; g[1] = j; g[2] = i; g[3] = j + 1; g[4] = i + 1; g[5] = j + 1; }
  }
  __device__ __forceinline__ void triangle(int t, double* out) const {
    int g[6];
    grid(t, g);
    for (int c = 0; c < 3; c++) {
      const int i = g[2 * c], j = g[2 * c + 1];
      out[3 * c] = x0 + i * spacing;
      out[3 * c + 1] = at(i, j);
      out[3 * c + 2] = z0 + j * spacing;
    }
  }
};

// src/mesh.js:67-71 (flags)
__device__ __forceinline__ int edge_flags(double bend, double length) {
  if (bend != bend) return FRONT | BACK;
  const double flat = length * MESH_FLAT;
  return (bend < -flat ? FRONT : 0) | (bend > flat ? BACK : 0);
}

// src/mesh.js:81-95 (fold)
__device__ __forceinline__ void edge_fold(const double* p, const double* n, int e, double fx, double fy, double fz,
                                          double* out, int o) {
  const int a = 3 * e, b = 3 * ((e + 1) % 3), c = 3 * ((e + 2) % 3);
  double ex = p[b] - p[a], ey = p[b + 1] - p[a + 1], ez = p[b + 2] - p[a + 2];
  const double el = dsqrt(ex * ex + ey * ey + ez * ez);
  ex /= el; ey /= el; ez /= el;
  double mx = ey * n[2] - ez * n[1], my = ez * n[0] - ex * n[2], mz = ex * n[1] - ey * n[0];
  if (mx * (p[c] - p[a]) + my * (p[c + 1] - p[a + 1]) + mz * (p[c + 2] - p[a + 2]) > 0) { mx = -mx; my = -my; mz = -mz; }
  double dx = fx - p[a], dy = fy - p[a + 1], dz = fz - p[a + 2];
  const double along = dx * ex + dy * ey + dz * ez;
  dx -= along * ex; dy -= along * ey; dz -= along * ez;
  const double dl = dsqrt(dx * dx + dy * dy + dz * dz);
  out[o] = (dx * mx + dy * my + dz * mz) / dl;
  out[o + 1] = (dx * n[0] + dy * n[1] + dz * n[2]) / dl;
}

// src/mesh.js:104-113 (normal)
__device__ __forceinline__ double triangle_normal(const double* p, double* out) {
  const double ux = p[3] - p[0], uy = p[4] - p[1], uz = p[5] - p[2];
  const double vx = p[6] - p[0], vy = p[7] - p[1], vz = p[8] - p[2];
  const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  const double l = dsqrt(nx * nx + ny * ny + nz * nz);
  if (l > 0) { out[0] = nx / l; out[1] = ny / l; out[2] = nz / l; }
  return l;
}

// src/mesh.js:409-455 (Heightfield.edge): triangle t's edge flags, and its folds.
__device__ int field_edge(const Field& F, int t, double* folds) {
  int G[6], H[6];
  double tri[9], n3[3] = {0, 0, 0};
  F.grid(t, G);
  F.triangle(t, tri);
  triangle_normal(tri, n3);
  const int cells = F.columns - 1;
  int out = 0;
  for (int e = 0; e < 3; e++) {
    const int ai = G[2 * e], aj = G[2 * e + 1], bi = G[(2 * e + 2) % 6], bj = G[(2 * e + 3) % 6];
    double bend = __builtin_nan("");
    folds[2 * e] = 0; folds[2 * e + 1] = 0;
    const int cell = t >> 1, ci = cell % cells, cj = (cell - ci) / cells;
    int ni = ci, nj = cj, other = -1;
    if (ai != bi && aj != bj) {
      other = t ^ 1;
    } else {
      if (ai == bi) ni = ai == ci ? ci - 1 : ci + 1;
      else nj = aj == cj ? cj - 1 : cj + 1;
      if (ni >= 0 && nj >= 0 && ni < cells && nj < F.rows - 1) {
        for (int k = 0; k < 2; k++) {
          const int u = 2 * (nj * cells + ni) + k;
          F.grid(u, H);
          int has = 0;
          for (int c = 0; c < 3; c++)
            if ((H[2 * c] == ai && H[2 * c + 1] == aj) || (H[2 * c] == bi && H[2 * c + 1] == bj)) has++;
          if (has == 2) other = u;
        }
      }
    }
    if (other >= 0) {
      F.grid(other, H);
      for (int c = 0; c < 3; c++) {
        const int fi = H[2 * c], fj = H[2 * c + 1];
        if ((fi == ai && fj == aj) || (fi == bi && fj == bj)) continue;
        const double X = F.x0 + fi * F.spacing, Y = F.at(fi, fj), Z = F.z0 + fj * F.spacing;
        const double fx = X - tri[0], fy = Y - tri[1], fz = Z - tri[2];
        bend = fx * n3[0] + fy * n3[1] + fz * n3[2];
        edge_fold(tri, n3, e, X, Y, Z, folds, 2 * e);
      }
    }
    const double ex = (bi - ai) * F.spacing, ey = F.at(bi, bj) - F.at(ai, aj), ez = (bj - aj) * F.spacing;
    const int f = edge_flags(bend, dsqrt(ex * ex + ey * ey + ez * ez));
    out |= ((f & FRONT) ? FRONT << e : 0) | ((f & BACK) ? BACK << e : 0);
  }
  return out | ONE_SIDED;
}

// What one triangle gives: up to MOST points.
struct Touch {
  int count;
  double normal[3];
  double points[3 * LEVEL_MOST];
  double separations[LEVEL_MOST];
  uint32_t ids[LEVEL_MOST];
};

__device__ __forceinline__ void touch_add(Touch& out, double x, double y, double z, double separation, uint32_t id) {
  if (out.count == LEVEL_MOST) return;
  const int i = out.count++;
  out.points[3 * i] = x; out.points[3 * i + 1] = y; out.points[3 * i + 2] = z;
  out.separations[i] = separation;
  out.ids[i] = id;
}
__device__ __forceinline__ void touch_aim(Touch& out, double x, double y, double z) {
  out.normal[0] = x; out.normal[1] = y; out.normal[2] = z;
}

// src/triangle.js:527-551 (clip)
__device__ __forceinline__ int level_clip(const double* s, const uint32_t* sid, int n, double* d, uint32_t* did,
                                          double px, double py, double pz, double mx, double my, double mz, int plane) {
  int k = 0;
  for (int i = 0; i < n; i++) {
    const int j = i + 1 < n ? i + 1 : 0;
    const double fi = (s[3 * i] - px) * mx + (s[3 * i + 1] - py) * my + (s[3 * i + 2] - pz) * mz;
    const double fj = (s[3 * j] - px) * mx + (s[3 * j + 1] - py) * my + (s[3 * j + 2] - pz) * mz;
    if (fi >= 0) {
      d[3 * k] = s[3 * i]; d[3 * k + 1] = s[3 * i + 1]; d[3 * k + 2] = s[3 * i + 2]; did[k++] = sid[i];
    }
    if ((fi >= 0) != (fj >= 0)) {
      const double u = fi / (fi - fj);
      d[3 * k] = s[3 * i] + u * (s[3 * j] - s[3 * i]);
      d[3 * k + 1] = s[3 * i + 1] + u * (s[3 * j + 1] - s[3 * i + 1]);
      d[3 * k + 2] = s[3 * i + 2] + u * (s[3 * j + 2] - s[3 * i + 2]);
      did[k++] = 256 + 16 * plane + i;
    }
  }
  return k;
}

// src/triangle.js's module state for one touch(): the triangle turned to face the shape,
// its seams and folds, and the clip buffers (ping-pong: `which` says which holds the source).
struct Toucher {
  double up[3], P[9], ahead[3], folds[6];
  int corners, behind;
  bool one_sided;
  uint32_t mix;
  double lo, hi;
  double EV[9], EL[3], Q[3];
  double seg[6];
  int cut[2];
  double buf[2][3 * LEVEL_CLIP];
  uint32_t bid[2][LEVEL_CLIP];
  int which;
  int face_k, face_side, incident;
  Placed H;

  __device__ __forceinline__ double* src() { return buf[which]; }
  __device__ __forceinline__ uint32_t* src_id() { return bid[which]; }
  __device__ __forceinline__ int clip_into(int n, double px, double py, double pz, double mx, double my, double mz, int plane) {
    const int k = level_clip(buf[which], bid[which], n, buf[which ^ 1], bid[which ^ 1], px, py, pz, mx, my, mz, plane);
    which ^= 1;
    return k;
  }

  // src/triangle.js:129-134 (seam)
  __device__ __forceinline__ bool seam(int region) const {
    if (region == 6) return false;
    if (region >= 3) return ((corners >> (region - 3)) & 1) == 0;
    return ((corners >> region) & 1) == 0 && ((corners >> ((region + 2) % 3)) & 1) == 0;
  }

  // src/triangle.js:162-190 (beside)
  __device__ bool beside(double x, double y, double z, double nx, double ny, double nz, double tolerance) const {
    const double t2 = tolerance * tolerance;
    int near = 0;
    bool back = false;
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, b = 3 * ((e + 1) % 3), c = 3 * ((e + 2) % 3);
      const double ex = P[b] - P[a], ey = P[b + 1] - P[a + 1], ez = P[b + 2] - P[a + 2];
      const double px = x - P[a], py = y - P[a + 1], pz = z - P[a + 2];
      const double ee = ex * ex + ey * ey + ez * ez, f = js_max(0, js_min(1, (px * ex + py * ey + pz * ez) / ee));
      const double qx = px - f * ex, qy = py - f * ey, qz = pz - f * ez;
      if (qx * qx + qy * qy + qz * qz > t2) continue;
      near++;
      double mx = ey * ahead[2] - ez * ahead[1], my = ez * ahead[0] - ex * ahead[2], mz = ex * ahead[1] - ey * ahead[0];
      if (mx * (P[c] - P[a]) + my * (P[c + 1] - P[a + 1]) + mz * (P[c + 2] - P[a + 2]) > 0) { mx = -mx; my = -my; mz = -mz; }
      const double l = dsqrt(mx * mx + my * my + mz * mz), out = (nx * mx + ny * my + nz * mz) / l;
      if (out < -MESH_FLAT) back = true;
      const bool seamed = ((corners >> e) & 1) == 0;
      const double gm = seamed ? 1 : folds[2 * e], gn = seamed ? 0 : folds[2 * e + 1];
      if (gm * out + gn * (nx * ahead[0] + ny * ahead[1] + nz * ahead[2]) > MESH_FLAT) return true;
    }
    return near == 1 && back;
  }

  // src/triangle.js:200-219 (wedge)
  __device__ bool wedge(double x, double y, double z) const {
    int past = -1;
    double mx = 0, my = 0, mz = 0, l = 0;
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, b = 3 * ((e + 1) % 3), c = 3 * ((e + 2) % 3);
      const double ex = P[b] - P[a], ey = P[b + 1] - P[a + 1], ez = P[b + 2] - P[a + 2];
      double ox = ey * ahead[2] - ez * ahead[1], oy = ez * ahead[0] - ex * ahead[2], oz = ex * ahead[1] - ey * ahead[0];
      if (ox * (P[c] - P[a]) + oy * (P[c + 1] - P[a + 1]) + oz * (P[c + 2] - P[a + 2]) > 0) { ox = -ox; oy = -oy; oz = -oz; }
      if ((x - P[a]) * ox + (y - P[a + 1]) * oy + (z - P[a + 2]) * oz <= 0) continue;
      if (past >= 0) return false;
      past = e; mx = ox; my = oy; mz = oz; l = dsqrt(ex * ex + ey * ey + ez * ez);
    }
    if (past < 0 || ((corners >> past) & 1) != 0) return false;
    const int a = 3 * past;
    const double gm = folds[2 * past] / l, gn = folds[2 * past + 1];
    const double gx = gm * mx + gn * ahead[0], gy = gm * my + gn * ahead[1], gz = gm * mz + gn * ahead[2];
    return (x - P[a]) * gx + (y - P[a + 1]) * gy + (z - P[a + 2]) * gz < 0;
  }

  // src/triangle.js:221-237 (region)
  __device__ int region(double x, double y, double z, double tolerance) const {
    const double t2 = tolerance * tolerance;
    for (int v = 0; v < 3; v++) {
      const double dx = x - P[3 * v], dy = y - P[3 * v + 1], dz = z - P[3 * v + 2];
      if (dx * dx + dy * dy + dz * dz <= t2) return v;
    }
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, b = 3 * ((e + 1) % 3);
      const double ex = P[b] - P[a], ey = P[b + 1] - P[a + 1], ez = P[b + 2] - P[a + 2];
      const double px = x - P[a], py = y - P[a + 1], pz = z - P[a + 2];
      const double ee = ex * ex + ey * ey + ez * ez, f = (px * ex + py * ey + pz * ez) / ee;
      const double qx = px - f * ex, qy = py - f * ey, qz = pz - f * ez;
      if (f >= 0 && f <= 1 && qx * qx + qy * qy + qz * qz <= t2) return 3 + e;
    }
    return 6;
  }

  __device__ __forceinline__ int at(double u, double v, int r, double ax, double ay, double az, double abx, double aby,
                                    double abz, double acx, double acy, double acz) {
    Q[0] = ax + u * abx + v * acx; Q[1] = ay + u * aby + v * acy; Q[2] = az + u * abz + v * acz;
    return r;
  }

  // src/triangle.js:240-270 (nearest, Ericson 5.1.5)
  __device__ int nearest(double x, double y, double z) {
    const double ax = P[0], ay = P[1], az = P[2];
    const double abx = P[3] - ax, aby = P[4] - ay, abz = P[5] - az, acx = P[6] - ax, acy = P[7] - ay, acz = P[8] - az;
    const double apx = x - ax, apy = y - ay, apz = z - az;
    const double d1 = abx * apx + aby * apy + abz * apz, d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= 0 && d2 <= 0) return at(0, 0, 0, ax, ay, az, abx, aby, abz, acx, acy, acz);
    const double bpx = x - P[3], bpy = y - P[4], bpz = z - P[5];
    const double d3 = abx * bpx + aby * bpy + abz * bpz, d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0 && d4 <= d3) return at(1, 0, 1, ax, ay, az, abx, aby, abz, acx, acy, acz);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return at(d1 / (d1 - d3), 0, 3, ax, ay, az, abx, aby, abz, acx, acy, acz);
    const double cpx = x - P[6], cpy = y - P[7], cpz = z - P[8];
    const double d5 = abx * cpx + aby * cpy + abz * cpz, d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0 && d5 <= d6) return at(0, 1, 2, ax, ay, az, abx, aby, abz, acx, acy, acz);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return at(0, d2 / (d2 - d6), 5, ax, ay, az, abx, aby, abz, acx, acy, acz);
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
      const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return at(1 - w, w, 4, ax, ay, az, abx, aby, abz, acx, acy, acz);
    }
    const double den = 1 / (va + vb + vc);
    return at(vb * den, vc * den, 6, ax, ay, az, abx, aby, abz, acx, acy, acz);
  }

  __device__ __forceinline__ double above(double x, double y, double z) const {
    return (x - P[0]) * up[0] + (y - P[1]) * up[1] + (z - P[2]) * up[2];
  }

  // src/triangle.js:273-276 (turn)
  __device__ __forceinline__ void turn() {
    up[0] = -up[0]; up[1] = -up[1]; up[2] = -up[2];
    const int c = corners;
    corners = behind;
    behind = c;
  }

  // src/triangle.js:292-334 (sphere)
  __device__ void sphere(const Shape& s, double margin, double tolerance, Touch& out) {
    const double* c = s.center;
    const double r = s.half[0];
    int where = nearest(c[0], c[1], c[2]);
    double nx = c[0] - Q[0], ny = c[1] - Q[1], nz = c[2] - Q[2];
    const double d = dsqrt(nx * nx + ny * ny + nz * nz);
    double sep;
    if (one_sided && above(c[0], c[1], c[2]) < 0) {
      if (where != 6) return;
      nx = up[0]; ny = up[1]; nz = up[2];
      sep = above(c[0], c[1], c[2]) - r;
    } else if (seam(where)) {
      const double h = above(c[0], c[1], c[2]);
      const double fx = c[0] - h * up[0] - Q[0], fy = c[1] - h * up[1] - Q[1], fz = c[2] - h * up[2] - Q[2];
      if (fx * fx + fy * fy + fz * fz > tolerance * tolerance) return;
      nx = up[0]; ny = up[1]; nz = up[2];
      sep = h - r;
      where = 6;
    } else if (d > r / 67108864) {
      nx /= d; ny /= d; nz /= d;
      if (beside(Q[0], Q[1], Q[2], nx, ny, nz, tolerance)) return;
      sep = d - r;
    } else {
      nx = up[0]; ny = up[1]; nz = up[2];
      sep = above(c[0], c[1], c[2]) - r;
    }
    if (sep > margin) return;
    touch_aim(out, -nx, -ny, -nz);
    const double m = r + sep / 2;
    touch_add(out, c[0] - nx * m, c[1] - ny * m, c[2] - nz * m, sep, (uint32_t)where);
  }

  // src/triangle.js:338-360 (overTriangle)
  __device__ bool over_triangle() {
    const double nx = up[0], ny = up[1], nz = up[2];
    cut[0] = -1; cut[1] = -1;
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, b = 3 * ((e + 1) % 3), o = 3 * ((e + 2) % 3);
      const double ex = P[b] - P[a], ey = P[b + 1] - P[a + 1], ez = P[b + 2] - P[a + 2];
      double mx = ny * ez - nz * ey, my = nz * ex - nx * ez, mz = nx * ey - ny * ex;
      const double cx = P[o] - P[a], cy = P[o + 1] - P[a + 1], cz = P[o + 2] - P[a + 2];
      if (mx * cx + my * cy + mz * cz < 0) { mx = -mx; my = -my; mz = -mz; }
      const double f0 = (seg[0] - P[a]) * mx + (seg[1] - P[a + 1]) * my + (seg[2] - P[a + 2]) * mz;
      const double f1 = (seg[3] - P[a]) * mx + (seg[4] - P[a + 1]) * my + (seg[5] - P[a + 2]) * mz;
      if (f0 < 0 && f1 < 0) return false;
      if (f0 < 0 || f1 < 0) {
        const double u = f0 / (f0 - f1);
        const int into = f0 < 0 ? 0 : 3;
        cut[into / 3] = e;
        seg[into] = seg[0] + u * (seg[3] - seg[0]);
        seg[into + 1] = seg[1] + u * (seg[4] - seg[1]);
        seg[into + 2] = seg[2] + u * (seg[5] - seg[2]);
      }
    }
    return true;
  }

  // src/triangle.js:362-414 (capsule)
  __device__ void capsule(const Shape& s, const Shape& t, double margin, double tolerance, Touch& out) {
    const double* c = s.center;
    const double* a = s.axes;
    const double r = s.half[0], h = s.half[1];
    double na[3], nb[3];
    const double d = gjk_closest(s, t, na, nb);
    double nx = 0, ny = 0, nz = 0, sep = 0;
    int where = 6;
    bool held = false;
    const bool apart = d > r / 67108864;
    if (apart) {
      nx = (na[0] - nb[0]) / d; ny = (na[1] - nb[1]) / d; nz = (na[2] - nb[2]) / d;
      sep = d - r;
      where = region(nb[0], nb[1], nb[2], tolerance);
      held = seam(where) || beside(nb[0], nb[1], nb[2], nx, ny, nz, tolerance)
          || (one_sided && nx * up[0] + ny * up[1] + nz * up[2] < 0);
    }
    seg[0] = c[0] - h * a[3]; seg[1] = c[1] - h * a[4]; seg[2] = c[2] - h * a[5];
    seg[3] = c[0] + h * a[3]; seg[4] = c[1] + h * a[4]; seg[5] = c[2] + h * a[5];
    const double e0 = above(seg[0], seg[1], seg[2]), e1 = above(seg[3], seg[4], seg[5]);
    const bool over = over_triangle();
    const double h0 = over ? above(seg[0], seg[1], seg[2]) : 0, h1 = over ? above(seg[3], seg[4], seg[5]) : 0;
    const double k = !one_sided && h0 < 0 && h1 < 0 ? -1 : 1;
    const double s0 = over ? k * h0 - r : __builtin_inf(), s1 = over ? k * h1 - r : __builtin_inf();
    const double face = js_min(s0, s1);
    const double whole = js_min(k * e0, k * e1) - r;
    if (over && (!apart || held || (face <= sep + tolerance && whole >= sep - tolerance))) {
      const double ux = k * up[0], uy = k * up[1], uz = k * up[2];
      touch_aim(out, -ux, -uy, -uz);
      for (int e = 0; e < 2; e++) {
        const double se = e == 0 ? s0 : s1;
        if (se > margin || (cut[e] >= 0 && ((corners >> cut[e]) & 1) == 0)) continue;
        const double m = r + se / 2;
        touch_add(out, seg[3 * e] - ux * m, seg[3 * e + 1] - uy * m, seg[3 * e + 2] - uz * m, se, 16 + e);
      }
      return;
    }
    if (!apart || held || sep > margin) return;
    touch_aim(out, -nx, -ny, -nz);
    const double m = r + sep / 2;
    touch_add(out, na[0] - nx * m, na[1] - ny * m, na[2] - nz * m, sep, (uint32_t)where);
  }

  // src/triangle.js:418-422 (reach)
  __device__ __forceinline__ static double box_reach(const Shape& b, double x, double y, double z) {
    const double* A = b.axes;
    const double* h = b.half;
    return dabs(A[0] * x + A[1] * y + A[2] * z) * h[0] + dabs(A[3] * x + A[4] * y + A[5] * z) * h[1]
        + dabs(A[6] * x + A[7] * y + A[8] * z) * h[2];
  }

  // src/triangle.js:436-443 (span)
  __device__ __forceinline__ void span(const double* c, double x, double y, double z) {
    double m = __builtin_inf(), M = -__builtin_inf();
    for (int v = 0; v < 3; v++) {
      const double d = (P[3 * v] - c[0]) * x + (P[3 * v + 1] - c[1]) * y + (P[3 * v + 2] - c[2]) * z;
      m = js_min(m, d); M = js_max(M, d);
    }
    lo = m; hi = M;
  }

  // src/triangle.js:451-459 (edges)
  __device__ __forceinline__ void edges() {
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, q = 3 * ((e + 1) % 3);
      const double vx = P[q] - P[a], vy = P[q + 1] - P[a + 1], vz = P[q + 2] - P[a + 2];
      EV[a] = vx; EV[a + 1] = vy; EV[a + 2] = vz;
      EL[e] = dsqrt(vx * vx + vy * vy + vz * vz);
    }
  }

  // src/triangle.js:616-620 (bottom)
  __device__ double bottom(int n) {
    const double* s = src();
    double m = __builtin_inf();
    for (int i = 0; i < n; i++) m = js_min(m, above(s[3 * i], s[3 * i + 1], s[3 * i + 2]));
    return m;
  }

  // src/triangle.js:623-650 (clipBox)
  __device__ int clip_box(const Shape& b) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    int k = 0;
    double most = 0;
    for (int j = 0; j < 3; j++) {
      const double d = A[3 * j] * up[0] + A[3 * j + 1] * up[1] + A[3 * j + 2] * up[2];
      if (dabs(d) > dabs(most)) { most = d; k = j; }
    }
    const int side = most > 0 ? -1 : 1, p = (k + 1) % 3, q = (k + 2) % 3;
    face_k = k; face_side = side;
    which = 0;
    double* s = src();
    uint32_t* sid = src_id();
    for (int i = 0; i < 4; i++) {
      const double cp = CORNER_P[i], cq = CORNER_Q[i];
      for (int x = 0; x < 3; x++)
        s[3 * i + x] = c[x] + side * h[k] * A[3 * k + x] + cp * h[p] * A[3 * p + x] + cq * h[q] * A[3 * q + x];
      sid[i] = i;
    }
    int n = 4;
    for (int e = 0; e < 3 && n > 0; e++) {
      const int a = 3 * e, q2 = 3 * ((e + 1) % 3), o = 3 * ((e + 2) % 3);
      const double vx = P[q2] - P[a], vy = P[q2 + 1] - P[a + 1], vz = P[q2 + 2] - P[a + 2];
      double mx = up[1] * vz - up[2] * vy, my = up[2] * vx - up[0] * vz, mz = up[0] * vy - up[1] * vx;
      if (mx * (P[o] - P[a]) + my * (P[o + 1] - P[a + 1]) + mz * (P[o + 2] - P[a + 2]) < 0) { mx = -mx; my = -my; mz = -mz; }
      n = clip_into(n, P[a], P[a + 1], P[a + 2], mx, my, mz, e);
    }
    return n;
  }

  // src/triangle.js:564-609 (triangleFace)
  __device__ bool triangle_face(const Shape& b, double margin, double tolerance, Touch& out) {
    const int n = clip_box(b);
    bool reached = false;
    for (int i = 0; i < n; i++) {
      const double* s = src();
      const double x = s[3 * i], y = s[3 * i + 1], z = s[3 * i + 2], sep = above(x, y, z);
      if (sep > margin) continue;
      reached = true;
      const uint32_t id = src_id()[i];
      if (id >= 4 && seam(region(x - up[0] * sep, y - up[1] * sep, z - up[2] * sep, tolerance))) continue;
      touch_add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, id < 4 ? 32 + id : mix + 32 + id);
    }
    {
      const double* c = b.center;
      const double* A = b.axes;
      const double* h = b.half;
      const int k = face_k, p = (k + 1) % 3, q = (k + 2) % 3;
      for (int i = 0; i < 4; i++) {
        const double cp = CORNER_P[i], cq = CORNER_Q[i];
        const double x = c[0] + face_side * h[k] * A[3 * k] + cp * h[p] * A[3 * p] + cq * h[q] * A[3 * q];
        const double y = c[1] + face_side * h[k] * A[3 * k + 1] + cp * h[p] * A[3 * p + 1] + cq * h[q] * A[3 * q + 1];
        const double z = c[2] + face_side * h[k] * A[3 * k + 2] + cp * h[p] * A[3 * p + 2] + cq * h[q] * A[3 * q + 2];
        const double sep = above(x, y, z);
        if (sep > margin || !wedge(x, y, z)) continue;
        reached = true;
        touch_add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, 32 + i);
      }
    }
    if (out.count > 0) touch_aim(out, -up[0], -up[1], -up[2]);
    return reached;
  }

  // src/triangle.js:653-679 (boxFace)
  __device__ bool box_face(const Shape& b, int k, int sign, double margin, double tolerance, Touch& out) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    const double ux = sign * A[3 * k], uy = sign * A[3 * k + 1], uz = sign * A[3 * k + 2];
    which = 0;
    for (int i = 0; i < 9; i++) buf[0][i] = P[i];
    bid[0][0] = 0; bid[0][1] = 1; bid[0][2] = 2;
    int n = 3;
    for (int plane = 0; plane < 4; plane++) {
      const int j = (k + 1 + (plane >> 1)) % 3;
      const double s = (plane & 1) ? -1 : 1;
      const double mx = -s * A[3 * j], my = -s * A[3 * j + 1], mz = -s * A[3 * j + 2];
      n = clip_into(n, c[0] + s * h[j] * A[3 * j], c[1] + s * h[j] * A[3 * j + 1], c[2] + s * h[j] * A[3 * j + 2], mx, my,
                    mz, plane);
      if (n == 0) return false;
    }
    for (int i = 0; i < n; i++) {
      const double* sp = src();
      const double x = sp[3 * i], y = sp[3 * i + 1], z = sp[3 * i + 2];
      const double sep = (x - c[0]) * ux + (y - c[1]) * uy + (z - c[2]) * uz - h[k];
      if (sep > margin || seam(region(x, y, z, tolerance)) || beside(x, y, z, -ux, -uy, -uz, tolerance)) continue;
      touch_add(out, x - ux * sep / 2, y - uy * sep / 2, z - uz * sep / 2, sep, mix + 64 + src_id()[i]);
    }
    if (out.count == 0) return false;
    touch_aim(out, ux, uy, uz);
    return true;
  }

  // src/triangle.js:682-715 (edgeContact)
  __device__ bool box_edge_contact(const Shape& b, int k, int e, double nx, double ny, double nz, double margin,
                                   double tolerance, Touch& out) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    double ox = c[0], oy = c[1], oz = c[2];
    for (int j = 0; j < 3; j++) {
      if (j == k) continue;
      const double s = A[3 * j] * nx + A[3 * j + 1] * ny + A[3 * j + 2] * nz >= 0 ? 1 : -1;
      ox += s * h[j] * A[3 * j]; oy += s * h[j] * A[3 * j + 1]; oz += s * h[j] * A[3 * j + 2];
    }
    const double ux = A[3 * k], uy = A[3 * k + 1], uz = A[3 * k + 2];
    const int a = 3 * e, q = 3 * ((e + 1) % 3);
    const double vx = P[q] - P[a], vy = P[q + 1] - P[a + 1], vz = P[q + 2] - P[a + 2];
    const double rx = ox - P[a], ry = oy - P[a + 1], rz = oz - P[a + 2];
    const double vv = vx * vx + vy * vy + vz * vz, uv = ux * vx + uy * vy + uz * vz;
    const double ur = ux * rx + uy * ry + uz * rz, vr = vx * rx + vy * ry + vz * rz;
    const double den = vv - uv * uv;
    double s = den > 0 ? (uv * vr - vv * ur) / den : 0;
    s = js_max(-h[k], js_min(h[k], s));
    double w = (vr + s * uv) / vv;
    w = js_max(0, js_min(1, w));
    s = js_max(-h[k], js_min(h[k], uv * w - ur));
    const double bx = ox + s * ux, by = oy + s * uy, bz = oz + s * uz;
    const double tx = P[a] + w * vx, ty = P[a + 1] + w * vy, tz = P[a + 2] + w * vz;
    if (seam(region(tx, ty, tz, tolerance)) || beside(tx, ty, tz, -nx, -ny, -nz, tolerance)) return false;
    const double sep = (tx - bx) * nx + (ty - by) * ny + (tz - bz) * nz;
    if (sep > margin) return false;
    touch_aim(out, nx, ny, nz);
    touch_add(out, (bx + tx) / 2, (by + ty) / 2, (bz + tz) / 2, sep, mix + 128 + 4 * k + e);
    return true;
  }

  // src/triangle.js:723-731 (closestPoints)
  __device__ void closest_points(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    double na[3], nb[3];
    const double d = gjk_closest(b, t, na, nb);
    if (!(d > 0) || d > margin) return;
    const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
    if (seam(region(nb[0], nb[1], nb[2], tolerance)) || beside(nb[0], nb[1], nb[2], -nx, -ny, -nz, tolerance)) return;
    touch_aim(out, nx, ny, nz);
    touch_add(out, na[0] + nx * d / 2, na[1] + ny * d / 2, na[2] + nz * d / 2, d, mix + 255);
  }

  // src/triangle.js:463-515 (box)
  __device__ void box(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    const double faceT = above(c[0], c[1], c[2]) - box_reach(b, up[0], up[1], up[2]);
    if (faceT > margin) return;
    double faceB = -__builtin_inf();
    int axisB = 0, signB = 1;
    for (int k = 0; k < 3; k++) {
      span(c, A[3 * k], A[3 * k + 1], A[3 * k + 2]);
      const double plus = lo - h[k], minus = -hi - h[k];
      const double sk = js_max(plus, minus);
      if (sk > margin) return;
      if (sk > faceB) { faceB = sk; axisB = k; signB = plus >= minus ? 1 : -1; }
    }
    edges();
    double edgeSep = -__builtin_inf(), ex = 0, ey = 0, ez = 0;
    int edgeK = -1, edgeE = -1;
    for (int k = 0; k < 3; k++) {
      const double ux = A[3 * k], uy = A[3 * k + 1], uz = A[3 * k + 2];
      for (int e = 0; e < 3; e++) {
        const double vx = EV[3 * e], vy = EV[3 * e + 1], vz = EV[3 * e + 2];
        double mx = uy * vz - uz * vy, my = uz * vx - ux * vz, mz = ux * vy - uy * vx;
        const double l = dsqrt(mx * mx + my * my + mz * mz);
        if (!(l > PARALLEL * EL[e])) continue;
        mx /= l; my /= l; mz /= l;
        span(c, mx, my, mz);
        const double r = box_reach(b, mx, my, mz), plus = lo - r, minus = -hi - r;
        if (minus > plus) { mx = -mx; my = -my; mz = -mz; }
        const double se = js_max(plus, minus);
        if (se > margin) return;
        if (se > edgeSep) { edgeSep = se; ex = mx; ey = my; ez = mz; edgeK = k; edgeE = e; }
      }
    }
    bool won = true;
    if (edgeK >= 0 && edgeSep > js_max(faceT, faceB) + tolerance) box_edge_contact(b, edgeK, edgeE, ex, ey, ez, margin, tolerance, out);
    else if (faceB > faceT + tolerance) box_face(b, axisB, signB, margin, tolerance, out);
    else won = false;
    if (won && out.count == 0 && !one_sided && bottom(clip_box(b)) + above(c[0], c[1], c[2]) < 0) turn();
    const bool reached = out.count == 0 && triangle_face(b, margin, tolerance, out);
    if (out.count == 0 && !reached && js_max(js_max(faceT, faceB), edgeSep) > 0) closest_points(b, t, margin, tolerance, out);
  }

  // src/triangle.js:743-753 (hullSpan)
  __device__ void hull_span(double nx, double ny, double nz) {
    double m = __builtin_inf(), M = -__builtin_inf();
    for (int i = 0; i < H.V; i++) {
      double v[3];
      H.vert(i, v);
      const double d = v[0] * nx + v[1] * ny + v[2] * nz;
      if (d < m) m = d;
      if (d > M) M = d;
    }
    lo = m; hi = M;
  }

  // src/triangle.js:755-757 (low)
  __device__ __forceinline__ double low(double x, double y, double z) const {
    return js_min(js_min(P[0] * x + P[1] * y + P[2] * z, P[3] * x + P[4] * y + P[5] * z), P[6] * x + P[7] * y + P[8] * z);
  }

  // src/triangle.js:881-924 (clipHull)
  __device__ int clip_hull() {
    int lowc = 0;
    double lowest = __builtin_inf();
    for (int i = 0; i < H.V; i++) {
      double v[3];
      H.vert(i, v);
      const double d = v[0] * up[0] + v[1] * up[1] + v[2] * up[2];
      if (d < lowest) { lowest = d; lowc = i; }
    }
    int f = 0;
    double least = __builtin_inf();
    for (int k = 0; k < H.F; k++) {
      double pl[4];
      H.plane(k, pl);
      const double d = pl[0] * up[0] + pl[1] * up[1] + pl[2] * up[2];
      if (!(d < least)) continue;
      bool at_corner = false;
      for (int j = H.start[k]; j < H.start[k + 1] && !at_corner; j++) at_corner = H.loops[j] == lowc;
      if (at_corner) { least = d; f = k; }
    }
    incident = f;
    const int i0 = H.start[f], m = H.start[f + 1] - i0;
    which = 0;
    for (int k = 0; k < m; k++) {
      H.vert(H.loops[i0 + k], &buf[0][3 * k]);
      bid[0][k] = CORNER_ID + (uint32_t)H.loops[i0 + k];
    }
    int n = m;
    for (int e = 0; e < 3 && n > 0; e++) {
      const int a = 3 * e, q2 = 3 * ((e + 1) % 3), o = 3 * ((e + 2) % 3);
      const double vx = P[q2] - P[a], vy = P[q2 + 1] - P[a + 1], vz = P[q2 + 2] - P[a + 2];
      double mx = up[1] * vz - up[2] * vy, my = up[2] * vx - up[0] * vz, mz = up[0] * vy - up[1] * vx;
      if (mx * (P[o] - P[a]) + my * (P[o + 1] - P[a + 1]) + mz * (P[o + 2] - P[a + 2]) < 0) { mx = -mx; my = -my; mz = -mz; }
      n = clip_into(n, P[a], P[a + 1], P[a + 2], mx, my, mz, e);
    }
    return n;
  }

  // src/triangle.js:853-878 (triangleFaceHull)
  __device__ bool triangle_face_hull(double margin, double tolerance, Touch& out) {
    const int n = clip_hull();
    bool reached = false;
    for (int k = 0; k < n; k++) {
      const double* s = src();
      const double x = s[3 * k], y = s[3 * k + 1], z = s[3 * k + 2], sep = above(x, y, z);
      if (sep > margin) continue;
      reached = true;
      const uint32_t id = src_id()[k];
      const bool own = id >= CORNER_ID;
      if (!own && seam(region(x - up[0] * sep, y - up[1] * sep, z - up[2] * sep, tolerance))) continue;
      touch_add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, own ? id : mix + 32 + id);
    }
    for (int j = H.start[incident]; j < H.start[incident + 1]; j++) {
      double v[3];
      H.vert(H.loops[j], v);
      const double x = v[0], y = v[1], z = v[2], sep = above(x, y, z);
      if (sep > margin || !wedge(x, y, z)) continue;
      reached = true;
      touch_add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, CORNER_ID + (uint32_t)H.loops[j]);
    }
    if (out.count > 0) touch_aim(out, -up[0], -up[1], -up[2]);
    return reached;
  }

  // src/triangle.js:927-951 (hullFace)
  __device__ void hull_face_touch(int f, double margin, double tolerance, Touch& out) {
    double pl[4];
    H.plane(f, pl);
    const double nx = pl[0], ny = pl[1], nz = pl[2], d = pl[3];
    const int i0 = H.start[f], m = H.start[f + 1] - i0;
    which = 0;
    for (int i = 0; i < 9; i++) buf[0][i] = P[i];
    bid[0][0] = 0; bid[0][1] = 1; bid[0][2] = 2;
    int n = 3;
    for (int k = 0; k < m && n > 0; k++) {
      double vp[3], vq[3];
      H.vert(H.loops[i0 + k], vp);
      H.vert(H.loops[i0 + (k + 1 < m ? k + 1 : 0)], vq);
      const double wx = vq[0] - vp[0], wy = vq[1] - vp[1], wz = vq[2] - vp[2];
      const double mx = ny * wz - nz * wy, my = nz * wx - nx * wz, mz = nx * wy - ny * wx;
      n = clip_into(n, vp[0], vp[1], vp[2], mx, my, mz, k & 15);
    }
    for (int k = 0; k < n; k++) {
      const double* s = src();
      const double x = s[3 * k], y = s[3 * k + 1], z = s[3 * k + 2];
      const double sep = x * nx + y * ny + z * nz - d;
      if (sep > margin || seam(region(x, y, z, tolerance)) || beside(x, y, z, -nx, -ny, -nz, tolerance)) continue;
      touch_add(out, x - nx * sep / 2, y - ny * sep / 2, z - nz * sep / 2, sep, mix + 64 + src_id()[k]);
    }
    if (out.count > 0) touch_aim(out, nx, ny, nz);
  }

  // src/triangle.js:954-983 (hullEdge)
  __device__ void hull_edge(int k, int e, double nx, double ny, double nz, double margin, double tolerance, Touch& out) {
    double vi[3], vj[3];
    H.vert(H.edges[2 * k], vi);
    H.vert(H.edges[2 * k + 1], vj);
    const double ox = vi[0], oy = vi[1], oz = vi[2];
    const double ux = vj[0] - ox, uy = vj[1] - oy, uz = vj[2] - oz;
    const int a = 3 * e, q = 3 * ((e + 1) % 3);
    const double vx = P[q] - P[a], vy = P[q + 1] - P[a + 1], vz = P[q + 2] - P[a + 2];
    const double rx = ox - P[a], ry = oy - P[a + 1], rz = oz - P[a + 2];
    const double uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz, uv = ux * vx + uy * vy + uz * vz;
    const double ur = ux * rx + uy * ry + uz * rz, vr = vx * rx + vy * ry + vz * rz;
    const double den = uu * vv - uv * uv;
    double s = den > 0 ? js_max(0, js_min(1, (uv * vr - vv * ur) / den)) : 0;
    double w = (uv * s + vr) / vv;
    if (w < 0) { w = 0; s = js_max(0, js_min(1, -ur / uu)); }
    else if (w > 1) { w = 1; s = js_max(0, js_min(1, (uv - ur) / uu)); }
    const double bx = ox + s * ux, by = oy + s * uy, bz = oz + s * uz;
    const double tx = P[a] + w * vx, ty = P[a + 1] + w * vy, tz = P[a + 2] + w * vz;
    if (seam(region(tx, ty, tz, tolerance)) || beside(tx, ty, tz, -nx, -ny, -nz, tolerance)) return;
    const double sep = (tx - bx) * nx + (ty - by) * ny + (tz - bz) * nz;
    if (sep > margin) return;
    if (sep > 0 && (s == 0 || s == 1 || w == 0 || w == 1)) return;
    touch_aim(out, nx, ny, nz);
    touch_add(out, (bx + tx) / 2, (by + ty) / 2, (bz + tz) / 2, sep, mix + 384 + ((3 * k + e) & 127));
  }

  // src/triangle.js:767-826 (hull)
  __device__ void hull(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    H.set(b);
    const double faceT = H.lowest(up[0], up[1], up[2]) - (P[0] * up[0] + P[1] * up[1] + P[2] * up[2]);
    if (faceT > margin) return;
    double faceB = -__builtin_inf();
    int fb = 0;
    for (int f = 0; f < H.F; f++) {
      double pl[4];
      H.plane(f, pl);
      const double sb = low(pl[0], pl[1], pl[2]) - pl[3];
      if (sb > margin) return;
      if (sb > faceB) { faceB = sb; fb = f; }
    }
    edges();
    const double* c = H.c;
    const double loose = LOOSE * (dabs(c[0]) + dabs(c[1]) + dabs(c[2]) + H.reach);
    const double r = H.inner > loose ? H.inner - loose : -__builtin_inf(), best = js_max(faceT, faceB);
    double edgeSep = -__builtin_inf(), ex = 0, ey = 0, ez = 0;
    int edgeK = -1, edgeE = -1;
    for (int k = 0; k < H.E; k++) {
      double vi[3], vj[3];
      H.vert(H.edges[2 * k], vi);
      H.vert(H.edges[2 * k + 1], vj);
      const double ux = vj[0] - vi[0], uy = vj[1] - vi[1], uz = vj[2] - vi[2];
      const double lu = dsqrt(ux * ux + uy * uy + uz * uz);
      for (int te = 0; te < 3; te++) {
        const double wx = EV[3 * te], wy = EV[3 * te + 1], wz = EV[3 * te + 2];
        double mx = uy * wz - uz * wy, my = uz * wx - ux * wz, mz = ux * wy - uy * wx;
        const double l = dsqrt(mx * mx + my * my + mz * mz);
        if (!(l > PARALLEL * lu * EL[te])) continue;
        mx /= l; my /= l; mz /= l;
        const double down = low(mx, my, mz), rise = low(-mx, -my, -mz), mc = mx * c[0] + my * c[1] + mz * c[2];
        if (js_max(down - mc, rise + mc) - r < best) continue;
        hull_span(mx, my, mz);
        const double plus = down - hi, minus = rise + lo;
        double top = hi;
        if (minus > plus) { mx = -mx; my = -my; mz = -mz; top = -lo; }
        if (top - (mx * vi[0] + my * vi[1] + mz * vi[2]) > tolerance) continue;
        const double sep = js_max(plus, minus);
        if (sep > margin) return;
        if (sep > edgeSep) { edgeSep = sep; ex = mx; ey = my; ez = mz; edgeK = k; edgeE = te; }
      }
    }
    bool won = true;
    if (edgeK >= 0 && edgeSep > js_max(faceT, faceB) + tolerance) hull_edge(edgeK, edgeE, ex, ey, ez, margin, tolerance, out);
    else if (faceB > faceT + tolerance) hull_face_touch(fb, margin, tolerance, out);
    else won = false;
    if (won && out.count == 0 && !one_sided && bottom(clip_hull()) + above(b.center[0], b.center[1], b.center[2]) < 0)
      turn();
    const bool reached = out.count == 0 && triangle_face_hull(margin, tolerance, out);
    if (out.count == 0 && !reached && js_max(js_max(faceT, faceB), edgeSep) > 0) closest_points(b, t, margin, tolerance, out);
  }

  // src/triangle.js:104-124 (touch)
  __device__ int touch(const Shape& s, const Shape& t, int flags, double margin, double tolerance, Touch& out, int tri) {
    out.count = 0;
    mix = (uint32_t)(tri + 1) * 512u;
    for (int i = 0; i < 9; i++) P[i] = t.points[i];
    one_sided = (flags & ONE_SIDED) != 0;
    const double nx = t.axes[0], ny = t.axes[1], nz = t.axes[2];
    ahead[0] = nx; ahead[1] = ny; ahead[2] = nz;
    const bool front = one_sided || (s.center[0] - P[0]) * nx + (s.center[1] - P[1]) * ny + (s.center[2] - P[2]) * nz >= 0;
    const double k = front ? 1 : -1;
    up[0] = k * nx; up[1] = k * ny; up[2] = k * nz;
    corners = front ? flags & 7 : (flags >> 3) & 7;
    behind = front ? (flags >> 3) & 7 : flags & 7;
    if (s.kind == SPHERE) sphere(s, margin, tolerance, out);
    else if (s.kind == CAPSULE) capsule(s, t, margin, tolerance, out);
    else if (s.kind == HULL) hull(s, t, margin, tolerance, out);
    else box(s, t, margin, tolerance, out);
    return out.count;
  }
};

// A lane's scratch for touchLevel: `cap` points found (x, y, z, separation; id, group) and
// gathered, and `cap / LEVEL_MOST` group normals.
struct LevelScratch {
  double *at, *gat, *normals;
  uint32_t *id, *in, *gid;
};

// src/triangle.js:1140-1179 (reduce): at most four of the m gathered points into out.
__device__ void level_reduce(const LevelScratch& W, int m, double tolerance, Manifold& out) {
  const double* X = W.gat;
  auto S = [&](int k) { return W.gat[4 * k + 3]; };
  auto keep = [&](int k) {
    if (out.count == MAX_POINTS) return;
    const int i = out.count++;
    out.points[3 * i] = X[4 * k]; out.points[3 * i + 1] = X[4 * k + 1]; out.points[3 * i + 2] = X[4 * k + 2];
    out.separations[i] = S(k);
    out.ids[i] = W.gid[k];
  };
  if (m <= MAX_POINTS) {
    for (int k = 0; k < m; k++) keep(k);
    return;
  }
  double deepest = S(0);
  for (int k = 1; k < m; k++) if (S(k) < deepest) deepest = S(k);
  const double pressed = js_max(deepest, 0), touching = pressed + tolerance;
  int a = 0;
  while (S(a) > pressed) a++;
  int b = -1;
  double far = -1;
  for (int pass = 0; pass < 3 && b < 0; pass++) {
    const double within = pass == 0 ? pressed : pass == 1 ? touching : __builtin_inf();
    for (int k = 0; k < m; k++) {
      if (k == a || S(k) > within) continue;
      const double x = X[4 * k] - X[4 * a], y = X[4 * k + 1] - X[4 * a + 1], z = X[4 * k + 2] - X[4 * a + 2];
      const double d = x * x + y * y + z * z;
      if (d > far) { far = d; b = k; }
    }
  }
  const double nx = out.normal[0], ny = out.normal[1], nz = out.normal[2];
  int c = -1, d = -1;
  double most = 0, least = 0;
  const double abx = X[4 * b] - X[4 * a], aby = X[4 * b + 1] - X[4 * a + 1], abz = X[4 * b + 2] - X[4 * a + 2];
  for (int pass = 0; pass < 3 && (c < 0 || d < 0); pass++) {
    const double within = pass == 0 ? pressed : pass == 1 ? touching : __builtin_inf();
    const bool left = c < 0, right = d < 0;
    for (int k = 0; k < m; k++) {
      if (k == a || k == b || S(k) > within) continue;
      const double x = X[4 * k] - X[4 * a], y = X[4 * k + 1] - X[4 * a + 1], z = X[4 * k + 2] - X[4 * a + 2];
      const double area = (aby * z - abz * y) * nx + (abz * x - abx * z) * ny + (abx * y - aby * x) * nz;
      if (left && area > most) { most = area; c = k; }
      if (right && area < least) { least = area; d = k; }
    }
  }
  keep(a);
  keep(b);
  if (c >= 0) keep(c);
  if (d >= 0) keep(d);
}
