// ORACLE — test infrastructure only. Level geometry for the CPU restatement: a shape against
// the triangles of a heightfield. Included into collide.cpp's unit (inside its anonymous
// namespace, after GJK), so it shares that file's Shape, Manifold, Placed and gjk_closest.
//
//   src/mesh.js:37-125, 271-455   FLAT, flags, fold, normal; the Heightfield
//   src/triangle.js:104-986       touch: sphere, capsule, box and hull against one triangle
//   src/triangle.js:988-1183      touchLevel: every triangle under a box, grouped by normal,
//                                 reduced to four points a direction
#pragma once

constexpr double MESH_FLAT = 1.0 / 1048576;         // 2^-20, src/mesh.js:40 (FLAT)
constexpr int FRONT = 1, BACK = 8, ONE_SIDED = 64;  // src/mesh.js:47, 60
constexpr int LEVEL_MOST = 7;                       // src/triangle.js:102 (MOST)
constexpr uint32_t CORNER_ID = 1u << 30;            // src/triangle.js:840 (CORNER)
constexpr double SAME = MESH_FLAT * MESH_FLAT / 2;  // src/triangle.js:986

// src/mesh.js:67-71 (flags)
int edge_flags(double bend, double length) {
  if (bend != bend) return FRONT | BACK;
  const double flat = length * MESH_FLAT;
  return (bend < -flat ? FRONT : 0) | (bend > flat ? BACK : 0);
}

// src/mesh.js:81-95 (fold): the way into the triangle across edge e, along this one's outward
// square to the edge (m) and its normal n, into out[o], out[o + 1].
void edge_fold(const double* p, const double* n, int e, double fx, double fy, double fz, double* out, int o) {
  const int a = 3 * e, b = 3 * ((e + 1) % 3), c = 3 * ((e + 2) % 3);
  double ex = p[b] - p[a], ey = p[b + 1] - p[a + 1], ez = p[b + 2] - p[a + 2];
  const double el = std::sqrt(ex * ex + ey * ey + ez * ez);
  ex /= el; ey /= el; ez /= el;
  double mx = ey * n[2] - ez * n[1], my = ez * n[0] - ex * n[2], mz = ex * n[1] - ey * n[0];
  if (mx * (p[c] - p[a]) + my * (p[c + 1] - p[a + 1]) + mz * (p[c + 2] - p[a + 2]) > 0) { mx = -mx; my = -my; mz = -mz; }
  double dx = fx - p[a], dy = fy - p[a + 1], dz = fz - p[a + 2];
  const double along = dx * ex + dy * ey + dz * ez;
  dx -= along * ex; dy -= along * ey; dz -= along * ez;
  const double dl = std::sqrt(dx * dx + dy * dy + dz * dz);
  out[o] = (dx * mx + dy * my + dz * mz) / dl;
  out[o + 1] = (dx * n[0] + dy * n[1] + dz * n[2]) / dl;
}

// src/mesh.js:104-113 (normal): a triangle's unit normal into out (left as it was for none),
// and twice its area.
double triangle_normal(const double* p, double* out) {
  const double ux = p[3] - p[0], uy = p[4] - p[1], uz = p[5] - p[2];
  const double vx = p[6] - p[0], vy = p[7] - p[1], vz = p[8] - p[2];
  const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (l > 0) { out[0] = nx / l; out[1] = ny / l; out[2] = nz / l; }
  return l;
}

// src/mesh.js:288-455 (Heightfield): rows of `columns` heights, `spacing` apart, centred on
// its body: x along columns, z along rows. Cell (i, j)'s two triangles share the diagonal
// from (i, j) to (i + 1, j + 1); both face +y.
struct Heightfield {
  std::vector<double> heights;
  int columns = 0, rows = 0;
  double spacing = 1, x0 = 0, z0 = 0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, reach = 0;

  void init(const double* h, int cols, int nrows, double sp) {
    heights.assign(h, h + (size_t)cols * nrows);
    columns = cols;
    rows = nrows;
    spacing = sp;
    x0 = -((cols - 1) * sp) / 2;
    z0 = -((nrows - 1) * sp) / 2;
    double l = INFINITY, u = -INFINITY;
    for (double v : heights) { l = js_min(l, v); u = js_max(u, v); }
    lo[0] = x0; lo[1] = l; lo[2] = z0;
    hi[0] = -x0; hi[1] = u; hi[2] = -z0;
    reach = std::sqrt(x0 * x0 + z0 * z0 + js_max(l * l, u * u));
  }
  double at(int i, int j) const { return heights[(size_t)j * columns + i]; }

  // src/mesh.js:317-321 (grid): triangle t's corners as grid points (i, j pairs).
  void grid(int t, int* g) const {
    const int cell = t >> 1, cells = columns - 1, i = cell % cells, j = (cell - i) / cells;
    if ((t & 1) == 0) { g[0] = i; g[1] = j; g[2] = i + 1; g[3] = j + 1; g[4] = i + 1; g[5] = j; }
    else { g[0] = i; g[1] = j; g[2] = i; g[3] = j + 1; g[4] = i + 1; g[5] = j + 1; }
  }

  // src/mesh.js:323-331 (triangle): its corners in the body's frame.
  void triangle(int t, double* out) const {
    int g[6];
    grid(t, g);
    for (int c = 0; c < 3; c++) {
      const int i = g[2 * c], j = g[2 * c + 1];
      out[3 * c] = x0 + i * spacing;
      out[3 * c + 1] = at(i, j);
      out[3 * c + 2] = z0 + j * spacing;
    }
  }

  // src/mesh.js:388-401 (triangles): both triangles of every cell under the box whose
  // heights reach it, row by row.
  template <class F>
  void triangles(double bx0, double by0, double bz0, double bx1, double by1, double bz1, F&& visit) const {
    const double s = spacing;
    const int cells = columns - 1;
    const double i0 = js_max(0, std::floor((bx0 - x0) / s)), i1 = js_min(cells - 1, std::floor((bx1 - x0) / s));
    const double j0 = js_max(0, std::floor((bz0 - z0) / s)), j1 = js_min(rows - 2, std::floor((bz1 - z0) / s));
    if (!(i0 <= i1) || !(j0 <= j1)) return;
    for (int j = (int)j0; j <= (int)j1; j++) {
      for (int i = (int)i0; i <= (int)i1; i++) {
        const double a = at(i, j), b = at(i + 1, j), c = at(i, j + 1), d = at(i + 1, j + 1);
        if (js_max(js_max(js_max(a, b), c), d) < by0 || js_min(js_min(js_min(a, b), c), d) > by1) continue;
        const int t = 2 * (j * cells + i);
        visit(t);
        visit(t + 1);
      }
    }
  }

  // src/mesh.js:409-455 (edge): triangle t's edge flags, each edge's triangle across found
  // in this cell or the next, its folds into `folds`.
  int edge(int t, double* folds) const {
    int G[6], H[6];
    double tri[9], n3[3] = {0, 0, 0};
    grid(t, G);
    triangle(t, tri);
    triangle_normal(tri, n3);
    const int cells = columns - 1;
    int out = 0;
    for (int e = 0; e < 3; e++) {
      const int ai = G[2 * e], aj = G[2 * e + 1], bi = G[(2 * e + 2) % 6], bj = G[(2 * e + 3) % 6];
      double bend = NAN;
      folds[2 * e] = 0; folds[2 * e + 1] = 0;
      const int cell = t >> 1, ci = cell % cells, cj = (cell - ci) / cells;
      int ni = ci, nj = cj, other = -1;
      if (ai != bi && aj != bj) {
        other = t ^ 1;
      } else {
        if (ai == bi) ni = ai == ci ? ci - 1 : ci + 1;
        else nj = aj == cj ? cj - 1 : cj + 1;
        if (ni >= 0 && nj >= 0 && ni < cells && nj < rows - 1) {
          for (int k = 0; k < 2; k++) {
            const int u = 2 * (nj * cells + ni) + k;
            grid(u, H);
            int has = 0;
            for (int c = 0; c < 3; c++)
              if ((H[2 * c] == ai && H[2 * c + 1] == aj) || (H[2 * c] == bi && H[2 * c + 1] == bj)) has++;
            if (has == 2) other = u;
          }
        }
      }
      if (other >= 0) {
        grid(other, H);
        for (int c = 0; c < 3; c++) {
          const int fi = H[2 * c], fj = H[2 * c + 1];
          if ((fi == ai && fj == aj) || (fi == bi && fj == bj)) continue;
          const double X = x0 + fi * spacing, Y = at(fi, fj), Z = z0 + fj * spacing;
          const double fx = X - tri[0], fy = Y - tri[1], fz = Z - tri[2];
          bend = fx * n3[0] + fy * n3[1] + fz * n3[2];
          edge_fold(tri, n3, e, X, Y, Z, folds, 2 * e);
        }
      }
      const double ex = (bi - ai) * spacing, ey = at(bi, bj) - at(ai, aj), ez = (bj - aj) * spacing;
      const int f = edge_flags(bend, std::sqrt(ex * ex + ey * ey + ez * ez));
      out |= ((f & FRONT) ? FRONT << e : 0) | ((f & BACK) ? BACK << e : 0);
    }
    return out | ONE_SIDED;
  }
};

// What one triangle gives (src/triangle.js:1004 `one`): up to MOST points.
struct Touch {
  int count;
  double normal[3];
  double points[3 * LEVEL_MOST];
  double separations[LEVEL_MOST];
  uint32_t ids[LEVEL_MOST];
};

// src/triangle.js's module state for one touch(): the triangle turned to face the shape, its
// corners, seams and folds, and the clip buffers.
struct Toucher {
  double up[3], P[9], ahead[3];
  int corners = 0, behind = 0;
  bool one_sided = false;
  const double* folds = nullptr;
  uint32_t mix = 0;
  double lo = 0, hi = 0;
  double EV[9], EL[3], Q[3];
  double seg[6];
  int cut[2];
  // Clip buffers: a box face's 4 corners or a triangle, clipped by up to 16 planes.
  double bufA[3 * 24], bufB[3 * 24];
  uint32_t idA[24], idB[24];
  double* src = bufA;
  double* dst = bufB;
  uint32_t* src_id = idA;
  uint32_t* dst_id = idB;
  int face_k = 0, face_side = 1, incident = 0;
  Placed H;

  void reset_buffers() { src = bufA; dst = bufB; src_id = idA; dst_id = idB; }
  void flip() {
    std::swap(src, dst);
    std::swap(src_id, dst_id);
  }

  // src/triangle.js:129-134 (seam)
  bool seam(int region) const {
    if (region == 6) return false;
    if (region >= 3) return ((corners >> (region - 3)) & 1) == 0;
    return ((corners >> region) & 1) == 0 && ((corners >> ((region + 2) % 3)) & 1) == 0;
  }

  // src/triangle.js:162-190 (beside)
  bool beside(double x, double y, double z, double nx, double ny, double nz, double tolerance) const {
    if (folds == nullptr) return false;
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
      const double l = std::sqrt(mx * mx + my * my + mz * mz), out = (nx * mx + ny * my + nz * mz) / l;
      if (out < -MESH_FLAT) back = true;
      const bool seamed = ((corners >> e) & 1) == 0;
      const double gm = seamed ? 1 : folds[2 * e], gn = seamed ? 0 : folds[2 * e + 1];
      if (gm * out + gn * (nx * ahead[0] + ny * ahead[1] + nz * ahead[2]) > MESH_FLAT) return true;
    }
    return near == 1 && back;
  }

  // src/triangle.js:200-219 (wedge)
  bool wedge(double x, double y, double z) const {
    if (folds == nullptr) return false;
    int past = -1;
    double mx = 0, my = 0, mz = 0, l = 0;
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, b = 3 * ((e + 1) % 3), c = 3 * ((e + 2) % 3);
      const double ex = P[b] - P[a], ey = P[b + 1] - P[a + 1], ez = P[b + 2] - P[a + 2];
      double ox = ey * ahead[2] - ez * ahead[1], oy = ez * ahead[0] - ex * ahead[2], oz = ex * ahead[1] - ey * ahead[0];
      if (ox * (P[c] - P[a]) + oy * (P[c + 1] - P[a + 1]) + oz * (P[c + 2] - P[a + 2]) > 0) { ox = -ox; oy = -oy; oz = -oz; }
      if ((x - P[a]) * ox + (y - P[a + 1]) * oy + (z - P[a + 2]) * oz <= 0) continue;
      if (past >= 0) return false;
      past = e; mx = ox; my = oy; mz = oz; l = std::sqrt(ex * ex + ey * ey + ez * ez);
    }
    if (past < 0 || ((corners >> past) & 1) != 0) return false;
    const int a = 3 * past;
    const double gm = folds[2 * past] / l, gn = folds[2 * past + 1];
    const double gx = gm * mx + gn * ahead[0], gy = gm * my + gn * ahead[1], gz = gm * mz + gn * ahead[2];
    return (x - P[a]) * gx + (y - P[a + 1]) * gy + (z - P[a + 2]) * gz < 0;
  }

  // src/triangle.js:221-237 (region): 0-2 corners, 3-5 edges, 6 the face.
  int region(double x, double y, double z, double tolerance) const {
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

  // src/triangle.js:240-270 (nearest): the closest point on the triangle into Q (Ericson
  // 5.1.5), and its region.
  int nearest(double x, double y, double z) {
    const double ax = P[0], ay = P[1], az = P[2];
    const double abx = P[3] - ax, aby = P[4] - ay, abz = P[5] - az, acx = P[6] - ax, acy = P[7] - ay, acz = P[8] - az;
    const double apx = x - ax, apy = y - ay, apz = z - az;
    const double d1 = abx * apx + aby * apy + abz * apz, d2 = acx * apx + acy * apy + acz * apz;
    auto at = [&](double u, double v, int r) {
      Q[0] = ax + u * abx + v * acx; Q[1] = ay + u * aby + v * acy; Q[2] = az + u * abz + v * acz;
      return r;
    };
    if (d1 <= 0 && d2 <= 0) return at(0, 0, 0);
    const double bpx = x - P[3], bpy = y - P[4], bpz = z - P[5];
    const double d3 = abx * bpx + aby * bpy + abz * bpz, d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0 && d4 <= d3) return at(1, 0, 1);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return at(d1 / (d1 - d3), 0, 3);
    const double cpx = x - P[6], cpy = y - P[7], cpz = z - P[8];
    const double d5 = abx * cpx + aby * cpy + abz * cpz, d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0 && d5 <= d6) return at(0, 1, 2);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return at(0, d2 / (d2 - d6), 5);
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
      const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return at(1 - w, w, 4);
    }
    const double den = 1 / (va + vb + vc);
    return at(vb * den, vc * den, 6);
  }

  double above(double x, double y, double z) const {
    return (x - P[0]) * up[0] + (y - P[1]) * up[1] + (z - P[2]) * up[2];
  }

  // src/triangle.js:273-276 (turn)
  void turn() {
    up[0] = -up[0]; up[1] = -up[1]; up[2] = -up[2];
    std::swap(corners, behind);
  }

  static void add(Touch& out, double x, double y, double z, double separation, uint32_t id) {
    if (out.count == LEVEL_MOST) return;
    const int i = out.count++;
    out.points[3 * i] = x; out.points[3 * i + 1] = y; out.points[3 * i + 2] = z;
    out.separations[i] = separation;
    out.ids[i] = id;
  }
  static void aim(Touch& out, double x, double y, double z) {
    out.normal[0] = x; out.normal[1] = y; out.normal[2] = z;
  }

  // src/triangle.js:292-334 (sphere)
  void sphere(const Shape& s, double margin, double tolerance, Touch& out) {
    const double* c = s.center;
    const double r = s.half[0];
    int where = nearest(c[0], c[1], c[2]);
    double nx = c[0] - Q[0], ny = c[1] - Q[1], nz = c[2] - Q[2];
    const double d = std::sqrt(nx * nx + ny * ny + nz * nz);
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
    aim(out, -nx, -ny, -nz);
    const double m = r + sep / 2;
    add(out, c[0] - nx * m, c[1] - ny * m, c[2] - nz * m, sep, (uint32_t)where);
  }

  // src/triangle.js:338-360 (overTriangle): seg clipped to the prism over the triangle.
  bool over_triangle() {
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
  void capsule(const Shape& s, const Shape& t, double margin, double tolerance, Touch& out) {
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
    const double s0 = over ? k * h0 - r : INFINITY, s1 = over ? k * h1 - r : INFINITY;
    const double face = js_min(s0, s1);
    const double whole = js_min(k * e0, k * e1) - r;
    if (over && (!apart || held || (face <= sep + tolerance && whole >= sep - tolerance))) {
      const double ux = k * up[0], uy = k * up[1], uz = k * up[2];
      aim(out, -ux, -uy, -uz);
      for (int e = 0; e < 2; e++) {
        const double se = e == 0 ? s0 : s1;
        if (se > margin || (cut[e] >= 0 && ((corners >> cut[e]) & 1) == 0)) continue;
        const double m = r + se / 2;
        add(out, seg[3 * e] - ux * m, seg[3 * e + 1] - uy * m, seg[3 * e + 2] - uz * m, se, 16 + e);
      }
      return;
    }
    if (!apart || held || sep > margin) return;
    aim(out, -nx, -ny, -nz);
    const double m = r + sep / 2;
    add(out, na[0] - nx * m, na[1] - ny * m, na[2] - nz * m, sep, (uint32_t)where);
  }

  // src/triangle.js:418-422 (reach): how far a box reaches from its centre along unit u.
  static double box_reach(const Shape& b, double x, double y, double z) {
    const double* A = b.axes;
    const double* h = b.half;
    return std::fabs(A[0] * x + A[1] * y + A[2] * z) * h[0] + std::fabs(A[3] * x + A[4] * y + A[5] * z) * h[1]
        + std::fabs(A[6] * x + A[7] * y + A[8] * z) * h[2];
  }

  // src/triangle.js:436-443 (span)
  void span(const double* c, double x, double y, double z) {
    double m = INFINITY, M = -INFINITY;
    for (int v = 0; v < 3; v++) {
      const double d = (P[3 * v] - c[0]) * x + (P[3 * v + 1] - c[1]) * y + (P[3 * v + 2] - c[2]) * z;
      m = js_min(m, d); M = js_max(M, d);
    }
    lo = m; hi = M;
  }

  // src/triangle.js:451-459 (edges)
  void edges() {
    for (int e = 0; e < 3; e++) {
      const int a = 3 * e, q = 3 * ((e + 1) % 3);
      const double vx = P[q] - P[a], vy = P[q + 1] - P[a + 1], vz = P[q + 2] - P[a + 2];
      EV[a] = vx; EV[a + 1] = vy; EV[a + 2] = vz;
      EL[e] = std::sqrt(vx * vx + vy * vy + vz * vz);
    }
  }

  // src/triangle.js:463-515 (box)
  void box(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    const double faceT = above(c[0], c[1], c[2]) - box_reach(b, up[0], up[1], up[2]);
    if (faceT > margin) return;
    double faceB = -INFINITY;
    int axisB = 0, signB = 1;
    for (int k = 0; k < 3; k++) {
      span(c, A[3 * k], A[3 * k + 1], A[3 * k + 2]);
      const double plus = lo - h[k], minus = -hi - h[k];
      const double sk = js_max(plus, minus);
      if (sk > margin) return;
      if (sk > faceB) { faceB = sk; axisB = k; signB = plus >= minus ? 1 : -1; }
    }
    edges();
    double edgeSep = -INFINITY, ex = 0, ey = 0, ez = 0;
    int edgeK = -1, edgeE = -1;
    for (int k = 0; k < 3; k++) {
      const double ux = A[3 * k], uy = A[3 * k + 1], uz = A[3 * k + 2];
      for (int e = 0; e < 3; e++) {
        const double vx = EV[3 * e], vy = EV[3 * e + 1], vz = EV[3 * e + 2];
        double mx = uy * vz - uz * vy, my = uz * vx - ux * vz, mz = ux * vy - uy * vx;
        const double l = std::sqrt(mx * mx + my * my + mz * mz);
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

  // src/triangle.js:527-551 (clip): src to (x - p)·m >= 0 into dst; new points are named by
  // the plane that made them and the edge they are on.
  static int clip(const double* s, const uint32_t* sid, int n, double* d, uint32_t* did, double px, double py,
                  double pz, double mx, double my, double mz, int plane) {
    int k = 0;
    for (int i = 0; i < n; i++) {
      const int j = (i + 1) % n;
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

  // src/triangle.js:564-609 (triangleFace): the box's face most against the triangle,
  // clipped to it; whether any point came within the margin.
  bool triangle_face(const Shape& b, double margin, double tolerance, Touch& out) {
    static const double CP[4] = {-1, 1, 1, -1}, CQ[4] = {-1, -1, 1, 1};
    const int n = clip_box(b);
    bool reached = false;
    for (int i = 0; i < n; i++) {
      const double x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2], sep = above(x, y, z);
      if (sep > margin) continue;
      reached = true;
      const uint32_t id = src_id[i];
      if (id >= 4 && seam(region(x - up[0] * sep, y - up[1] * sep, z - up[2] * sep, tolerance))) continue;
      add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, id < 4 ? 32 + id : mix + 32 + id);
    }
    if (folds != nullptr) {
      const double* c = b.center;
      const double* A = b.axes;
      const double* h = b.half;
      const int k = face_k, p = (k + 1) % 3, q = (k + 2) % 3;
      for (int i = 0; i < 4; i++) {
        const double x = c[0] + face_side * h[k] * A[3 * k] + CP[i] * h[p] * A[3 * p] + CQ[i] * h[q] * A[3 * q];
        const double y = c[1] + face_side * h[k] * A[3 * k + 1] + CP[i] * h[p] * A[3 * p + 1] + CQ[i] * h[q] * A[3 * q + 1];
        const double z = c[2] + face_side * h[k] * A[3 * k + 2] + CP[i] * h[p] * A[3 * p + 2] + CQ[i] * h[q] * A[3 * q + 2];
        const double sep = above(x, y, z);
        if (sep > margin || !wedge(x, y, z)) continue;
        reached = true;
        add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, 32 + i);
      }
    }
    if (out.count > 0) aim(out, -up[0], -up[1], -up[2]);
    return reached;
  }

  // src/triangle.js:616-620 (bottom)
  double bottom(int n) const {
    double m = INFINITY;
    for (int i = 0; i < n; i++) m = js_min(m, above(src[3 * i], src[3 * i + 1], src[3 * i + 2]));
    return m;
  }

  // src/triangle.js:623-650 (clipBox)
  int clip_box(const Shape& b) {
    static const double CP[4] = {-1, 1, 1, -1}, CQ[4] = {-1, -1, 1, 1};
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    int k = 0;
    double most = 0;
    for (int j = 0; j < 3; j++) {
      const double d = A[3 * j] * up[0] + A[3 * j + 1] * up[1] + A[3 * j + 2] * up[2];
      if (std::fabs(d) > std::fabs(most)) { most = d; k = j; }
    }
    const int side = most > 0 ? -1 : 1, p = (k + 1) % 3, q = (k + 2) % 3;
    face_k = k; face_side = side;
    reset_buffers();
    for (int i = 0; i < 4; i++) {
      for (int x = 0; x < 3; x++)
        src[3 * i + x] = c[x] + side * h[k] * A[3 * k + x] + CP[i] * h[p] * A[3 * p + x] + CQ[i] * h[q] * A[3 * q + x];
      src_id[i] = i;
    }
    int n = 4;
    for (int e = 0; e < 3 && n > 0; e++) {
      const int a = 3 * e, q2 = 3 * ((e + 1) % 3), o = 3 * ((e + 2) % 3);
      const double vx = P[q2] - P[a], vy = P[q2 + 1] - P[a + 1], vz = P[q2 + 2] - P[a + 2];
      double mx = up[1] * vz - up[2] * vy, my = up[2] * vx - up[0] * vz, mz = up[0] * vy - up[1] * vx;
      if (mx * (P[o] - P[a]) + my * (P[o + 1] - P[a + 1]) + mz * (P[o + 2] - P[a + 2]) < 0) { mx = -mx; my = -my; mz = -mz; }
      n = clip(src, src_id, n, dst, dst_id, P[a], P[a + 1], P[a + 2], mx, my, mz, e);
      flip();
    }
    return n;
  }

  // src/triangle.js:653-679 (boxFace): the triangle clipped to box face k.
  bool box_face(const Shape& b, int k, int sign, double margin, double tolerance, Touch& out) {
    const double* c = b.center;
    const double* A = b.axes;
    const double* h = b.half;
    const double ux = sign * A[3 * k], uy = sign * A[3 * k + 1], uz = sign * A[3 * k + 2];
    reset_buffers();
    for (int i = 0; i < 9; i++) src[i] = P[i];
    src_id[0] = 0; src_id[1] = 1; src_id[2] = 2;
    int n = 3;
    for (int plane = 0; plane < 4; plane++) {
      const int j = (k + 1 + (plane >> 1)) % 3;
      const double s = (plane & 1) ? -1 : 1;
      const double mx = -s * A[3 * j], my = -s * A[3 * j + 1], mz = -s * A[3 * j + 2];
      n = clip(src, src_id, n, dst, dst_id, c[0] + s * h[j] * A[3 * j], c[1] + s * h[j] * A[3 * j + 1],
               c[2] + s * h[j] * A[3 * j + 2], mx, my, mz, plane);
      flip();
      if (n == 0) return false;
    }
    for (int i = 0; i < n; i++) {
      const double x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
      const double sep = (x - c[0]) * ux + (y - c[1]) * uy + (z - c[2]) * uz - h[k];
      if (sep > margin || seam(region(x, y, z, tolerance)) || beside(x, y, z, -ux, -uy, -uz, tolerance)) continue;
      add(out, x - ux * sep / 2, y - uy * sep / 2, z - uz * sep / 2, sep, mix + 64 + src_id[i]);
    }
    if (out.count == 0) return false;
    aim(out, ux, uy, uz);
    return true;
  }

  // src/triangle.js:682-715 (edgeContact): box edge along axis k against triangle edge e.
  bool box_edge_contact(const Shape& b, int k, int e, double nx, double ny, double nz, double margin, double tolerance,
                        Touch& out) {
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
    aim(out, nx, ny, nz);
    add(out, (bx + tx) / 2, (by + ty) / 2, (bz + tz) / 2, sep, mix + 128 + 4 * k + e);
    return true;
  }

  // src/triangle.js:723-731 (closestPoints): apart within the margin, GJK's closest points.
  void closest_points(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    double na[3], nb[3];
    const double d = gjk_closest(b, t, na, nb);
    if (!(d > 0) || d > margin) return;
    const double nx = (nb[0] - na[0]) / d, ny = (nb[1] - na[1]) / d, nz = (nb[2] - na[2]) / d;
    if (seam(region(nb[0], nb[1], nb[2], tolerance)) || beside(nb[0], nb[1], nb[2], -nx, -ny, -nz, tolerance)) return;
    aim(out, nx, ny, nz);
    add(out, na[0] + nx * d / 2, na[1] + ny * d / 2, na[2] + nz * d / 2, d, mix + 255);
  }

  // src/triangle.js:743-753 (hullSpan)
  void hull_span(double nx, double ny, double nz) {
    const double* v = H.v.data();
    const int V = H.hull->V;
    double m = INFINITY, M = -INFINITY;
    for (int i = 0; i < V; i++) {
      const double d = v[3 * i] * nx + v[3 * i + 1] * ny + v[3 * i + 2] * nz;
      if (d < m) m = d;
      if (d > M) M = d;
    }
    lo = m; hi = M;
  }

  // src/triangle.js:755-757 (low)
  double low(double x, double y, double z) const {
    return js_min(js_min(P[0] * x + P[1] * y + P[2] * z, P[3] * x + P[4] * y + P[5] * z), P[6] * x + P[7] * y + P[8] * z);
  }

  // src/triangle.js:767-826 (hull)
  void hull(const Shape& b, const Shape& t, double margin, double tolerance, Touch& out) {
    H.set(b);
    const double* v = H.v.data();
    const double* pl = H.planes.data();
    const Hull& h = *H.hull;
    const int F = h.F;
    const double faceT = H.lowest(up[0], up[1], up[2]) - (P[0] * up[0] + P[1] * up[1] + P[2] * up[2]);
    if (faceT > margin) return;
    double faceB = -INFINITY;
    int fb = 0;
    for (int f = 0; f < F; f++) {
      const double sb = low(pl[4 * f], pl[4 * f + 1], pl[4 * f + 2]) - pl[4 * f + 3];
      if (sb > margin) return;
      if (sb > faceB) { faceB = sb; fb = f; }
    }
    const int* e = h.edges.data();
    const int E = h.E;
    edges();
    const double* c = H.center;
    const double loose = LOOSE * (std::fabs(c[0]) + std::fabs(c[1]) + std::fabs(c[2]) + H.reach);
    const double r = H.inner > loose ? H.inner - loose : -INFINITY, best = js_max(faceT, faceB);
    double edgeSep = -INFINITY, ex = 0, ey = 0, ez = 0;
    int edgeK = -1, edgeE = -1;
    for (int k = 0; k < E; k++) {
      const int i = 3 * e[2 * k], j = 3 * e[2 * k + 1];
      const double ux = v[j] - v[i], uy = v[j + 1] - v[i + 1], uz = v[j + 2] - v[i + 2];
      const double lu = std::sqrt(ux * ux + uy * uy + uz * uz);
      for (int te = 0; te < 3; te++) {
        const double wx = EV[3 * te], wy = EV[3 * te + 1], wz = EV[3 * te + 2];
        double mx = uy * wz - uz * wy, my = uz * wx - ux * wz, mz = ux * wy - uy * wx;
        const double l = std::sqrt(mx * mx + my * my + mz * mz);
        if (!(l > PARALLEL * lu * EL[te])) continue;
        mx /= l; my /= l; mz /= l;
        const double down = low(mx, my, mz), rise = low(-mx, -my, -mz), mc = mx * c[0] + my * c[1] + mz * c[2];
        if (js_max(down - mc, rise + mc) - r < best) continue;
        hull_span(mx, my, mz);
        const double plus = down - hi, minus = rise + lo;
        double top = hi;
        if (minus > plus) { mx = -mx; my = -my; mz = -mz; top = -lo; }
        if (top - (mx * v[i] + my * v[i + 1] + mz * v[i + 2]) > tolerance) continue;
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

  // src/triangle.js:853-878 (triangleFaceHull)
  bool triangle_face_hull(double margin, double tolerance, Touch& out) {
    const int n = clip_hull();
    bool reached = false;
    for (int k = 0; k < n; k++) {
      const double x = src[3 * k], y = src[3 * k + 1], z = src[3 * k + 2], sep = above(x, y, z);
      if (sep > margin) continue;
      reached = true;
      const uint32_t id = src_id[k];
      const bool own = id >= CORNER_ID;
      if (!own && seam(region(x - up[0] * sep, y - up[1] * sep, z - up[2] * sep, tolerance))) continue;
      add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, own ? id : mix + 32 + id);
    }
    if (folds != nullptr) {
      const Hull& h = *H.hull;
      const double* v = H.v.data();
      for (int j = h.start[incident]; j < h.start[incident + 1]; j++) {
        const int i = 3 * h.loops[j];
        const double x = v[i], y = v[i + 1], z = v[i + 2], sep = above(x, y, z);
        if (sep > margin || !wedge(x, y, z)) continue;
        reached = true;
        add(out, x - up[0] * sep / 2, y - up[1] * sep / 2, z - up[2] * sep / 2, sep, CORNER_ID + (uint32_t)h.loops[j]);
      }
    }
    if (out.count > 0) aim(out, -up[0], -up[1], -up[2]);
    return reached;
  }

  // src/triangle.js:881-924 (clipHull): the incident face at the deepest corner, clipped.
  int clip_hull() {
    const Hull& h = *H.hull;
    const double* pl = H.planes.data();
    const double* v = H.v.data();
    const int F = h.F, V = h.V;
    int lowc = 0;
    double lowest = INFINITY;
    for (int i = 0; i < V; i++) {
      const double d = v[3 * i] * up[0] + v[3 * i + 1] * up[1] + v[3 * i + 2] * up[2];
      if (d < lowest) { lowest = d; lowc = i; }
    }
    int f = 0;
    double least = INFINITY;
    for (int k = 0; k < F; k++) {
      const double d = pl[4 * k] * up[0] + pl[4 * k + 1] * up[1] + pl[4 * k + 2] * up[2];
      if (!(d < least)) continue;
      bool at = false;
      for (int j = h.start[k]; j < h.start[k + 1] && !at; j++) at = h.loops[j] == lowc;
      if (at) { least = d; f = k; }
    }
    incident = f;
    const int i0 = h.start[f], m = h.start[f + 1] - i0;
    reset_buffers();
    for (int k = 0; k < m; k++) {
      const int i = 3 * h.loops[i0 + k];
      src[3 * k] = v[i]; src[3 * k + 1] = v[i + 1]; src[3 * k + 2] = v[i + 2];
      src_id[k] = CORNER_ID + (uint32_t)h.loops[i0 + k];
    }
    int n = m;
    for (int e = 0; e < 3 && n > 0; e++) {
      const int a = 3 * e, q2 = 3 * ((e + 1) % 3), o = 3 * ((e + 2) % 3);
      const double vx = P[q2] - P[a], vy = P[q2 + 1] - P[a + 1], vz = P[q2 + 2] - P[a + 2];
      double mx = up[1] * vz - up[2] * vy, my = up[2] * vx - up[0] * vz, mz = up[0] * vy - up[1] * vx;
      if (mx * (P[o] - P[a]) + my * (P[o + 1] - P[a + 1]) + mz * (P[o + 2] - P[a + 2]) < 0) { mx = -mx; my = -my; mz = -mz; }
      n = clip(src, src_id, n, dst, dst_id, P[a], P[a + 1], P[a + 2], mx, my, mz, e);
      flip();
    }
    return n;
  }

  // src/triangle.js:927-951 (hullFace): the triangle clipped to the sides of hull face f.
  void hull_face_touch(int f, double margin, double tolerance, Touch& out) {
    const Hull& h = *H.hull;
    const double* pl = H.planes.data();
    const double* v = H.v.data();
    const double nx = pl[4 * f], ny = pl[4 * f + 1], nz = pl[4 * f + 2], d = pl[4 * f + 3];
    const int i0 = h.start[f], m = h.start[f + 1] - i0;
    reset_buffers();
    for (int i = 0; i < 9; i++) src[i] = P[i];
    src_id[0] = 0; src_id[1] = 1; src_id[2] = 2;
    int n = 3;
    for (int k = 0; k < m && n > 0; k++) {
      const int p = 3 * h.loops[i0 + k], q = 3 * h.loops[i0 + (k + 1 < m ? k + 1 : 0)];
      const double wx = v[q] - v[p], wy = v[q + 1] - v[p + 1], wz = v[q + 2] - v[p + 2];
      const double mx = ny * wz - nz * wy, my = nz * wx - nx * wz, mz = nx * wy - ny * wx;
      n = clip(src, src_id, n, dst, dst_id, v[p], v[p + 1], v[p + 2], mx, my, mz, k & 15);
      flip();
    }
    for (int k = 0; k < n; k++) {
      const double x = src[3 * k], y = src[3 * k + 1], z = src[3 * k + 2];
      const double sep = x * nx + y * ny + z * nz - d;
      if (sep > margin || seam(region(x, y, z, tolerance)) || beside(x, y, z, -nx, -ny, -nz, tolerance)) continue;
      add(out, x - nx * sep / 2, y - ny * sep / 2, z - nz * sep / 2, sep, mix + 64 + src_id[k]);
    }
    if (out.count > 0) aim(out, nx, ny, nz);
  }

  // src/triangle.js:954-983 (hullEdge)
  void hull_edge(int k, int e, double nx, double ny, double nz, double margin, double tolerance, Touch& out) {
    const double* v = H.v.data();
    const int* ed = H.hull->edges.data();
    const int i = 3 * ed[2 * k], j = 3 * ed[2 * k + 1];
    const double ox = v[i], oy = v[i + 1], oz = v[i + 2];
    const double ux = v[j] - ox, uy = v[j + 1] - oy, uz = v[j + 2] - oz;
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
    aim(out, nx, ny, nz);
    add(out, (bx + tx) / 2, (by + ty) / 2, (bz + tz) / 2, sep, mix + 384 + ((3 * k + e) & 127));
  }

  // src/triangle.js:104-124 (touch): shape s against triangle t (world corners in t.points,
  // unit normal in t.axes[0..2]), its normal from s towards the triangle.
  int touch(const Shape& s, const Shape& t, int flags, const double* fold, double margin, double tolerance, Touch& out,
            int tri) {
    out.count = 0;
    mix = (uint32_t)(tri + 1) * 512u;
    for (int i = 0; i < 9; i++) P[i] = t.points[i];
    folds = fold;
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

// src/triangle.js:988-1183 (touchLevel, visit, reduce): where shape s touches heightfield g
// whose body is at c (3) turned by axes A (9), looking in the box (x0..z1, the field's frame):
// one manifold for each direction it is pushed in, normals from s towards the field.
struct LevelToucher {
  Toucher T;
  Touch one;
  std::vector<double> found_at, found_sep, normals;
  std::vector<uint32_t> found_id;
  std::vector<int> found_in;
  std::vector<double> gathered, gathered_sep;
  std::vector<uint32_t> gathered_id;
  int found = 0, groups = 0;
  double tolerance = 0;

  void visit(const Heightfield& g, const double* c, const double* A, const Shape& s, int t, double margin) {
    double local[9], folds[6];
    Shape tri{};
    tri.kind = TRIANGLE_KIND;
    tri.axes[0] = 1; tri.axes[4] = 1; tri.axes[8] = 1;
    g.triangle(t, local);
    double* p = tri.points;
    for (int v = 0; v < 3; v++) {
      const double x = local[3 * v], y = local[3 * v + 1], z = local[3 * v + 2];
      p[3 * v] = c[0] + A[0] * x + A[3] * y + A[6] * z;
      p[3 * v + 1] = c[1] + A[1] * x + A[4] * y + A[7] * z;
      p[3 * v + 2] = c[2] + A[2] * x + A[5] * y + A[8] * z;
    }
    tri.center[0] = (p[0] + p[3] + p[6]) / 3;
    tri.center[1] = (p[1] + p[4] + p[7]) / 3;
    tri.center[2] = (p[2] + p[5] + p[8]) / 3;
    triangle_normal(p, tri.axes);
    const int flags = g.edge(t, folds);
    if (T.touch(s, tri, flags, folds, margin, tolerance, one, t) == 0) return;
    const double nx = one.normal[0], ny = one.normal[1], nz = one.normal[2];
    int k = 0;
    while (k < groups && normals[3 * k] * nx + normals[3 * k + 1] * ny + normals[3 * k + 2] * nz < 1 - SAME) k++;
    if (k == groups) {
      normals.resize(3 * (groups + 1));
      normals[3 * k] = nx; normals[3 * k + 1] = ny; normals[3 * k + 2] = nz;
      groups++;
    }
    const double t2 = tolerance * tolerance;
    for (int q = 0; q < one.count; q++) {
      const double x = one.points[3 * q], y = one.points[3 * q + 1], z = one.points[3 * q + 2], sep = one.separations[q];
      int same = -1;
      for (int i = 0; i < found && same < 0; i++) {
        if (found_in[i] != k) continue;
        const double dx = found_at[3 * i] - x, dy = found_at[3 * i + 1] - y, dz = found_at[3 * i + 2] - z;
        if (dx * dx + dy * dy + dz * dz <= t2) same = i;
      }
      if (same >= 0) {
        if (sep < found_sep[same]) {
          found_at[3 * same] = x; found_at[3 * same + 1] = y; found_at[3 * same + 2] = z; found_sep[same] = sep;
        }
        if (one.ids[q] < 512 && found_id[same] >= 512) found_id[same] = one.ids[q];
        continue;
      }
      found_at.resize(3 * (found + 1)); found_sep.resize(found + 1); found_id.resize(found + 1); found_in.resize(found + 1);
      found_at[3 * found] = x; found_at[3 * found + 1] = y; found_at[3 * found + 2] = z;
      found_sep[found] = sep; found_id[found] = one.ids[q]; found_in[found] = k;
      found++;
    }
  }

  // src/triangle.js:1140-1179 (reduce): at most four of the m gathered points.
  void reduce(int m, Manifold& out) {
    const double* X = gathered.data();
    const double* S = gathered_sep.data();
    auto keep = [&](int k) {
      if (out.count == MAX_POINTS) return;
      const int i = out.count++;
      out.points[3 * i] = X[3 * k]; out.points[3 * i + 1] = X[3 * k + 1]; out.points[3 * i + 2] = X[3 * k + 2];
      out.separations[i] = S[k];
      out.ids[i] = gathered_id[k];
    };
    if (m <= MAX_POINTS) {
      for (int k = 0; k < m; k++) keep(k);
      return;
    }
    double deepest = S[0];
    for (int k = 1; k < m; k++) if (S[k] < deepest) deepest = S[k];
    const double pressed = js_max(deepest, 0), touching = pressed + tolerance;
    int a = 0;
    while (S[a] > pressed) a++;
    int b = -1;
    double far = -1;
    for (int pass = 0; pass < 3 && b < 0; pass++) {
      const double within = pass == 0 ? pressed : pass == 1 ? touching : INFINITY;
      for (int k = 0; k < m; k++) {
        if (k == a || S[k] > within) continue;
        const double x = X[3 * k] - X[3 * a], y = X[3 * k + 1] - X[3 * a + 1], z = X[3 * k + 2] - X[3 * a + 2];
        const double d = x * x + y * y + z * z;
        if (d > far) { far = d; b = k; }
      }
    }
    const double nx = out.normal[0], ny = out.normal[1], nz = out.normal[2];
    int c = -1, d = -1;
    double most = 0, least = 0;
    const double abx = X[3 * b] - X[3 * a], aby = X[3 * b + 1] - X[3 * a + 1], abz = X[3 * b + 2] - X[3 * a + 2];
    for (int pass = 0; pass < 3 && (c < 0 || d < 0); pass++) {
      const double within = pass == 0 ? pressed : pass == 1 ? touching : INFINITY;
      const bool left = c < 0, right = d < 0;
      for (int k = 0; k < m; k++) {
        if (k == a || k == b || S[k] > within) continue;
        const double x = X[3 * k] - X[3 * a], y = X[3 * k + 1] - X[3 * a + 1], z = X[3 * k + 2] - X[3 * a + 2];
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

  // src/triangle.js:1078-1131 (touchLevel); the manifolds into `out`, returns their count.
  int touch_level(const Heightfield& g, const double* c, const double* A, const Shape& s, const double* box,
                  double margin, double tol, std::vector<Manifold>& out) {
    tolerance = tol;
    found = 0;
    groups = 0;
    g.triangles(box[0], box[1], box[2], box[3], box[4], box[5], [&](int t) { visit(g, c, A, s, t, margin); });
    out.clear();
    for (int k = 0; k < groups; k++) {
      int m = 0;
      gathered.clear(); gathered_sep.clear(); gathered_id.clear();
      for (int i = 0; i < found; i++) {
        if (found_in[i] != k) continue;
        gathered.resize(3 * (m + 1)); gathered_sep.resize(m + 1); gathered_id.resize(m + 1);
        int j = m++;
        for (; j > 0 && gathered_id[j - 1] > found_id[i]; j--) {
          gathered[3 * j] = gathered[3 * j - 3]; gathered[3 * j + 1] = gathered[3 * j - 2]; gathered[3 * j + 2] = gathered[3 * j - 1];
          gathered_sep[j] = gathered_sep[j - 1]; gathered_id[j] = gathered_id[j - 1];
        }
        gathered[3 * j] = found_at[3 * i]; gathered[3 * j + 1] = found_at[3 * i + 1]; gathered[3 * j + 2] = found_at[3 * i + 2];
        gathered_sep[j] = found_sep[i]; gathered_id[j] = found_id[i];
      }
      Manifold mf{};
      mf.count = 0;
      mf.normal[0] = normals[3 * k]; mf.normal[1] = normals[3 * k + 1]; mf.normal[2] = normals[3 * k + 2];
      reduce(m, mf);
      out.push_back(mf);
    }
    return (int)out.size();
  }
};
