// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg may load it, and only as the checker; the product
// never links it.
//
// A line-for-line CPU restatement of the reference's soft step: everything
// src/world.js step() does between colouring and _catch (src/world.js:961-1031),
// i.e. contact.js inertias / prepare / makeSoft / warmStart / solve /
// restitution / store and world.js _accelerate / _move with tumble / untwist /
// advance. Same operations, same order, doubles, no fused multiply-add, so the
// bits are the reference's. Each function cites the lines it follows.
//
// Pinned against the reference itself: tests/golden holds the reference's own
// state before inertias() and after store() for steps of five scenes, and the
// velocities after every call of one step (tests/test_oracle_soft.py).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "jsmath.h"

namespace oracle {
namespace {

constexpr int P = 4;      // points per manifold, src/contact.js:128
constexpr int J = 12;     // turning terms per point, src/contact.js:139
constexpr int FJ = 24;    // per manifold, two tangents
constexpr int DYNAMIC = 0;
constexpr int SPHERE_SHAPE = 1, CAPSULE_SHAPE = 2;
constexpr int COLOURS = 32, OVERFLOW = COLOURS, RESTING = OVERFLOW + 1;  // src/graph.js:67-71
constexpr double STIFFNESS = 1.0 / 8;       // src/world.js:115
constexpr double DAMPING_RATIO = 10;        // src/world.js:118
constexpr double SPIN_TURN = 3.141592653589793;  // SPIN, src/world.js:171
constexpr double SPIN_DISC = 2.0 / 3;       // SPIN, src/contact.js:146
constexpr double COMPLIANCE = 1.0 / 1048576;  // 2^-20, src/forest.js:67
constexpr double ROUNDED = 1.0 / 17592186044416.0;  // 2^-44, src/world.js:2809
constexpr double MATH_PI = 3.141592653589793;

struct Soft {
  double bias_rate = 0, mass_scale = 1, impulse_scale = 0;
};

// src/contact.js:1590-1597
Soft make_soft(double hertz, double zeta, double h) {
  const double omega = 2 * MATH_PI * hertz;
  const double a1 = 2 * zeta + h * omega, a2 = h * omega * a1, a3 = 1 / (1 + a2);
  Soft s;
  s.bias_rate = omega / a1;
  s.mass_scale = a2 * a3;
  s.impulse_scale = a3;
  return s;
}

// src/contact.js:1520-1549
void locked_inverse(const double* ax, const double* I, int lock, double* out) {
  static const int SYM[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  const double i0 = I[0], i1 = I[1], i2 = I[2];
  const double x0 = ax[0], x1 = ax[1], x2 = ax[2];
  const double y0 = ax[3], y1 = ax[4], y2 = ax[5];
  const double z0 = ax[6], z1 = ax[7], z2 = ax[8];
  double M6[6];
  M6[0] = x0 * x0 * i0 + y0 * y0 * i1 + z0 * z0 * i2;
  M6[1] = x0 * x1 * i0 + y0 * y1 * i1 + z0 * z1 * i2;
  M6[2] = x0 * x2 * i0 + y0 * y2 * i1 + z0 * z2 * i2;
  M6[3] = x1 * x1 * i0 + y1 * y1 * i1 + z1 * z1 * i2;
  M6[4] = x1 * x2 * i0 + y1 * y2 * i1 + z1 * z2 * i2;
  M6[5] = x2 * x2 * i0 + y2 * y2 * i1 + z2 * z2 * i2;
  for (int k = 0; k < 6; k++) out[k] = 0;
  int first = -1, second = -1;
  for (int k = 0; k < 3; k++) {
    if ((lock >> k) & 1) continue;
    if (first < 0) first = k; else second = k;
  }
  if (first < 0) return;
  const int ff = SYM[4 * first];
  if (second < 0) {
    if (M6[ff] > 0) out[ff] = 1 / M6[ff];
    return;
  }
  const int ss = SYM[4 * second], fs = SYM[3 * first + second];
  const double det = M6[ff] * M6[ss] - M6[fs] * M6[fs];
  if (!(det > 0)) return;
  out[ff] = M6[ss] / det; out[ss] = M6[ff] / det; out[fs] = -M6[fs] / det;
}

// src/contact.js:1493-1508 (inverseAt)
void inverse_at(const OracleSoft* w, int s, const double* ax, double* out) {
  if (w->type[s] != DYNAMIC) { for (int k = 0; k < 6; k++) out[k] = 0; return; }
  if (w->lock[s] != 0) { locked_inverse(ax, w->inertia + 3 * s, w->lock[s], out); return; }
  const double* inv = w->inv_inertia + 3 * s;
  const double i0 = inv[0], i1 = inv[1], i2 = inv[2];
  const double x0 = ax[0], x1 = ax[1], x2 = ax[2];
  const double y0 = ax[3], y1 = ax[4], y2 = ax[5];
  const double z0 = ax[6], z1 = ax[7], z2 = ax[8];
  out[0] = x0 * x0 * i0 + y0 * y0 * i1 + z0 * z0 * i2;
  out[1] = x0 * x1 * i0 + y0 * y1 * i1 + z0 * z1 * i2;
  out[2] = x0 * x2 * i0 + y0 * y2 * i1 + z0 * z2 * i2;
  out[3] = x1 * x1 * i0 + y1 * y1 * i1 + z1 * z1 * i2;
  out[4] = x1 * x2 * i0 + y1 * y2 * i1 + z1 * z2 * i2;
  out[5] = x2 * x2 * i0 + y2 * y2 * i1 + z2 * z2 * i2;
}

// src/contact.js:1302-1310
void inertias(OracleSoft* w) {
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (w->frozen[s] != 0) { for (int k = 0; k < 6; k++) w->inv_i[6 * s + k] = 0; }
    else inverse_at(w, s, w->axes + 9 * s, w->inv_i + 6 * s);
  }
}

// src/contact.js:1716-1722 (along), 1709-1713 (turning)
double along(const double* invI, int o, double x, double y, double z) {
  return invI[o] * x * x + invI[o + 3] * y * y + invI[o + 5] * z * z
      + 2 * (invI[o + 1] * x * y + invI[o + 2] * x * z + invI[o + 4] * y * z);
}
double turning(const double* invI, int a, int b, double x, double y, double z) {
  const double k = along(invI, a * 6, x, y, z) + along(invI, b * 6, x, y, z);
  return k > 0 ? 1 / k : 0;
}

// src/contact.js:1565-1581
double lever(double* jac, const double* invI, double ima, double imb, int a, int b, double rax, double ray,
             double raz, double rbx, double rby, double rbz, double dx, double dy, double dz) {
  const double cax = ray * dz - raz * dy, cay = raz * dx - rax * dz, caz = rax * dy - ray * dx;
  const double cbx = rby * dz - rbz * dy, cby = rbz * dx - rbx * dz, cbz = rbx * dy - rby * dx;
  const int oa = a * 6, ob = b * 6;
  const double iax = invI[oa] * cax + invI[oa + 1] * cay + invI[oa + 2] * caz;
  const double iay = invI[oa + 1] * cax + invI[oa + 3] * cay + invI[oa + 4] * caz;
  const double iaz = invI[oa + 2] * cax + invI[oa + 4] * cay + invI[oa + 5] * caz;
  const double ibx = invI[ob] * cbx + invI[ob + 1] * cby + invI[ob + 2] * cbz;
  const double iby = invI[ob + 1] * cbx + invI[ob + 3] * cby + invI[ob + 4] * cbz;
  const double ibz = invI[ob + 2] * cbx + invI[ob + 4] * cby + invI[ob + 5] * cbz;
  jac[0] = cax; jac[1] = cay; jac[2] = caz;
  jac[3] = cbx; jac[4] = cby; jac[5] = cbz;
  jac[6] = iax; jac[7] = iay; jac[8] = iaz;
  jac[9] = ibx; jac[10] = iby; jac[11] = ibz;
  const double k = ima + imb + cax * iax + cay * iay + caz * iaz + cbx * ibx + cby * iby + cbz * ibz;
  return k > 0 ? 1 / k : 0;
}

// src/contact.js:1199-1220 (core) for a whole body: the anchor in s's frame into
// out, its world offset at the start into v3.
void core(const OracleSoft* w, int s, double x, double y, double z, double* out, double* v3) {
  const int shape = w->shape[s];
  if (shape == SPHERE_SHAPE) {
    out[0] = 0; out[1] = 0; out[2] = 0;
    v3[0] = 0; v3[1] = 0; v3[2] = 0;
    return;
  }
  rotate(w->rotation + 4 * s, x, y, z, true, v3);
  if (shape != CAPSULE_SHAPE) {
    out[0] = v3[0]; out[1] = v3[1]; out[2] = v3[2];
    v3[0] = x; v3[1] = y; v3[2] = z;
    return;
  }
  const double half = (w->dims[s * 3 + 1] - 2 * w->dims[s * 3]) * 0.5;
  out[0] = 0; out[1] = v3[1] < -half ? -half : v3[1] > half ? half : v3[1]; out[2] = 0;
  rotate(w->rotation + 4 * s, 0, out[1], 0, false, v3);
}

// src/contact.js:1346-1445 (prepareRows)
void prepare(OracleSoft* w, int from, int to) {
  const double* invI = w->inv_i;
  const double *p = w->position, *v = w->velocity, *om = w->angular;
  double v3[3];
  for (int m = from; m < to; m++) {
    const int a = w->a[m], b = w->b[m], a3 = a * 3, b3 = b * 3;
    const double ima = w->inv_mass[a], imb = w->inv_mass[b];
    const int m3 = 3 * m;
    const double nx = w->normal[m3], ny = w->normal[m3 + 1], nz = w->normal[m3 + 2];
    const int count = w->points[m];
    if (count <= 2) {
      w->roll_mass1[m] = turning(invI, a, b, w->t1[m3], w->t1[m3 + 1], w->t1[m3 + 2]);
      w->roll_mass2[m] = turning(invI, a, b, w->t2[m3], w->t2[m3 + 1], w->t2[m3 + 2]);
    }
    w->spin_mass[m] = turning(invI, a, b, nx, ny, nz);
    const double tol = w->tolerance[m];
    double weight = 0, ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    double ex = 0, ey = 0, ez = 0, fx = 0, fy = 0, fz = 0;
    for (int k = 0; k < count; k++) {
      const int i = m * P + k, i3 = 3 * i;
      const double px = w->point[i3], py = w->point[i3 + 1], pz = w->point[i3 + 2];
      const double rax = px - p[a3], ray = py - p[a3 + 1], raz = pz - p[a3 + 2];
      const double rbx = px + w->shift[i3] - p[b3], rby = py + w->shift[i3 + 1] - p[b3 + 1],
                   rbz = pz + w->shift[i3 + 2] - p[b3 + 2];
      w->ra[i3] = rax; w->ra[i3 + 1] = ray; w->ra[i3 + 2] = raz;
      w->rb[i3] = rbx; w->rb[i3 + 1] = rby; w->rb[i3 + 2] = rbz;
      core(w, a, rax, ray, raz, w->la + i3, v3);
      const double cax = v3[0], cay = v3[1], caz = v3[2];
      core(w, b, rbx, rby, rbz, w->lb + i3, v3);
      w->adjusted[i] = w->separation[i] - ((v3[0] - cax) * nx + (v3[1] - cay) * ny + (v3[2] - caz) * nz);
      w->normal_mass[i] = lever(w->jac + J * i, invI, ima, imb, a, b, rax, ray, raz, rbx, rby, rbz, nx, ny, nz);
      w->max_normal_impulse[i] = 0;
      const double s = w->separation[i] / tol, near = s <= 1 ? 1 : s >= 2 ? 0 : 2 - s;
      weight += near;
      ax += near * rax; ay += near * ray; az += near * raz; bx += near * rbx; by += near * rby; bz += near * rbz;
      ex += rax; ey += ray; ez += raz; fx += rbx; fy += rby; fz += rbz;
      const double dvx = v[b3] + om[b3 + 1] * rbz - om[b3 + 2] * rby - v[a3] - om[a3 + 1] * raz + om[a3 + 2] * ray;
      const double dvy = v[b3 + 1] + om[b3 + 2] * rbx - om[b3] * rbz - v[a3 + 1] - om[a3 + 2] * rax + om[a3] * raz;
      const double dvz = v[b3 + 2] + om[b3] * rby - om[b3 + 1] * rbx - v[a3 + 2] - om[a3] * ray + om[a3 + 1] * rax;
      w->relative_velocity[i] = dvx * nx + dvy * ny + dvz * nz;
    }
    if (count >= 2) {
      const double* jac = w->jac;
      const int kb = m * P * P;
      const double s = ima + imb;
      for (int k = 0; k < count; k++) {
        const int jk = J * (m * P + k);
        for (int l = 0; l <= k; l++) {
          const int jl = J * (m * P + l);
          const double e = s + jac[jk] * jac[jl + 6] + jac[jk + 1] * jac[jl + 7] + jac[jk + 2] * jac[jl + 8]
              + jac[jk + 3] * jac[jl + 9] + jac[jk + 4] * jac[jl + 10] + jac[jk + 5] * jac[jl + 11];
          w->block[kb + P * k + l] = e; w->block[kb + P * l + k] = e;
        }
      }
    }
    if (weight == 0) { weight = count; ax = ex; ay = ey; az = ez; bx = fx; by = fy; bz = fz; }
    ax /= weight; ay /= weight; az /= weight; bx /= weight; by /= weight; bz /= weight;
    w->anchor_a[m3] = ax; w->anchor_a[m3 + 1] = ay; w->anchor_a[m3 + 2] = az;
    w->anchor_b[m3] = bx; w->anchor_b[m3 + 1] = by; w->anchor_b[m3 + 2] = bz;
    for (int k = 0; k < count; k++) {
      const int i3 = 3 * (m * P + k);
      double dx = w->ra[i3] - ax, dy = w->ra[i3 + 1] - ay, dz = w->ra[i3 + 2] - az;
      const double al = dx * nx + dy * ny + dz * nz;
      dx -= al * nx; dy -= al * ny; dz -= al * nz;
      w->lever[m * P + k] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    double* jac = w->fjac + FJ * m;
    lever(jac, invI, ima, imb, a, b, ax, ay, az, bx, by, bz, w->t1[m3], w->t1[m3 + 1], w->t1[m3 + 2]);
    lever(jac + 12, invI, ima, imb, a, b, ax, ay, az, bx, by, bz, w->t2[m3], w->t2[m3 + 1], w->t2[m3 + 2]);
    const double k11 = ima + imb + jac[0] * jac[6] + jac[1] * jac[7] + jac[2] * jac[8]
        + jac[3] * jac[9] + jac[4] * jac[10] + jac[5] * jac[11];
    const double k22 = ima + imb + jac[12] * jac[18] + jac[13] * jac[19] + jac[14] * jac[20]
        + jac[15] * jac[21] + jac[16] * jac[22] + jac[17] * jac[23];
    const double k12 = jac[12] * jac[6] + jac[13] * jac[7] + jac[14] * jac[8]
        + jac[15] * jac[9] + jac[16] * jac[10] + jac[17] * jac[11];
    const double det = k11 * k22 - k12 * k12;
    if (det > 0) {
      w->tangent_mass[m3] = k22 / det; w->tangent_mass[m3 + 1] = -k12 / det; w->tangent_mass[m3 + 2] = k11 / det;
    } else {
      w->tangent_mass[m3] = 0; w->tangent_mass[m3 + 1] = 0; w->tangent_mass[m3 + 2] = 0;
    }
  }
}

// src/contact.js:2096-2104
inline void write_back(double* v, double* om, double im, int s3, double vx, double vy, double vz, double wx,
                       double wy, double wz) {
  if (im == 0) return;
  v[s3] = vx; v[s3 + 1] = vy; v[s3 + 2] = vz; om[s3] = wx; om[s3 + 1] = wy; om[s3 + 2] = wz;
}

// src/contact.js:1616-1663 (warmRows)
void warm_rows(OracleSoft* w, int from, int to) {
  double *v = w->velocity, *om = w->angular;
  const double* invI = w->inv_i;
  const double* jac = w->jac;
  for (int m = from; m < to; m++) {
    const int a = w->a[m], b = w->b[m], a3 = a * 3, b3 = b * 3, m3 = 3 * m;
    const double ima = w->inv_mass[a], imb = w->inv_mass[b];
    const double nx = w->normal[m3], ny = w->normal[m3 + 1], nz = w->normal[m3 + 2];
    const double t1x = w->t1[m3], t1y = w->t1[m3 + 1], t1z = w->t1[m3 + 2];
    const double t2x = w->t2[m3], t2y = w->t2[m3 + 1], t2z = w->t2[m3 + 2];
    double vax = v[a3], vay = v[a3 + 1], vaz = v[a3 + 2], wax = om[a3], way = om[a3 + 1], waz = om[a3 + 2];
    double vbx = v[b3], vby = v[b3 + 1], vbz = v[b3 + 2], wbx = om[b3], wby = om[b3 + 1], wbz = om[b3 + 2];
    for (int k = 0; k < w->points[m]; k++) {
      const int i = m * P + k, j = J * i;
      const double ln = w->normal_impulse[i];
      const double px = nx * ln, py = ny * ln, pz = nz * ln;
      vax -= ima * px; vay -= ima * py; vaz -= ima * pz;
      vbx += imb * px; vby += imb * py; vbz += imb * pz;
      wax -= jac[j + 6] * ln; way -= jac[j + 7] * ln; waz -= jac[j + 8] * ln;
      wbx += jac[j + 9] * ln; wby += jac[j + 10] * ln; wbz += jac[j + 11] * ln;
    }
    {
      const double* fj = w->fjac + FJ * m;
      const double l1 = w->tangent_impulse1[m], l2 = w->tangent_impulse2[m];
      const double px = t1x * l1 + t2x * l2, py = t1y * l1 + t2y * l2, pz = t1z * l1 + t2z * l2;
      vax -= ima * px; vay -= ima * py; vaz -= ima * pz;
      vbx += imb * px; vby += imb * py; vbz += imb * pz;
      wax -= fj[6] * l1 + fj[18] * l2; way -= fj[7] * l1 + fj[19] * l2; waz -= fj[8] * l1 + fj[20] * l2;
      wbx += fj[9] * l1 + fj[21] * l2; wby += fj[10] * l1 + fj[22] * l2; wbz += fj[11] * l1 + fj[23] * l2;
    }
    {
      const double r1 = w->roll_impulse1[m], r2 = w->roll_impulse2[m], sp = w->spin_impulse[m];
      const int oa = a * 6, ob = b * 6;
      const double lx = t1x * r1 + t2x * r2 + nx * sp, ly = t1y * r1 + t2y * r2 + ny * sp,
                   lz = t1z * r1 + t2z * r2 + nz * sp;
      wax -= invI[oa] * lx + invI[oa + 1] * ly + invI[oa + 2] * lz;
      way -= invI[oa + 1] * lx + invI[oa + 3] * ly + invI[oa + 4] * lz;
      waz -= invI[oa + 2] * lx + invI[oa + 4] * ly + invI[oa + 5] * lz;
      wbx += invI[ob] * lx + invI[ob + 1] * ly + invI[ob + 2] * lz;
      wby += invI[ob + 1] * lx + invI[ob + 3] * ly + invI[ob + 4] * lz;
      wbz += invI[ob + 2] * lx + invI[ob + 4] * ly + invI[ob + 5] * lz;
    }
    write_back(v, om, ima, a3, vax, vay, vaz, wax, way, waz);
    write_back(v, om, imb, b3, vbx, vby, vbz, wbx, wby, wbz);
  }
}

// src/contact.js:1672-1724 (pushes): Gaussian elimination without pivoting.
void pushes(const double* K, int count, int set, const double* give, const double* wanted, double* solved) {
  if (set == 15) {
    const double g0 = 1 + give[0], g1 = 1 + give[1], g2 = 1 + give[2], g3 = 1 + give[3];
    double a00 = K[0] * g0, a01 = K[1], a02 = K[2], a03 = K[3];
    double a10 = K[4], a11 = K[5] * g1, a12 = K[6], a13 = K[7];
    double a20 = K[8], a21 = K[9], a22 = K[10] * g2, a23 = K[11];
    double a30 = K[12], a31 = K[13], a32 = K[14], a33 = K[15] * g3;
    double r0 = wanted[0], r1 = wanted[1], r2 = wanted[2], r3 = wanted[3], f;
    f = a10 / a00; a11 -= f * a01; a12 -= f * a02; a13 -= f * a03; r1 -= f * r0;
    f = a20 / a00; a21 -= f * a01; a22 -= f * a02; a23 -= f * a03; r2 -= f * r0;
    f = a30 / a00; a31 -= f * a01; a32 -= f * a02; a33 -= f * a03; r3 -= f * r0;
    f = a21 / a11; a22 -= f * a12; a23 -= f * a13; r2 -= f * r1;
    f = a31 / a11; a32 -= f * a12; a33 -= f * a13; r3 -= f * r1;
    f = a32 / a22; a33 -= f * a23; r3 -= f * r2;
    const double x3 = r3 / a33;
    const double x2 = (r2 - a23 * x3) / a22;
    const double x1 = (r1 - a12 * x2 - a13 * x3) / a11;
    const double x0 = (r0 - a01 * x1 - a02 * x2 - a03 * x3) / a00;
    solved[0] = x0; solved[1] = x1; solved[2] = x2; solved[3] = x3;
    (void)a10; (void)a20; (void)a30;
    return;
  }
  double AS[P * P], RS[P], XS[P];
  int IN[P];
  int n = 0;
  for (int k = 0; k < count; k++) {
    solved[k] = 0;
    if ((set >> k) & 1) IN[n++] = k;
  }
  for (int r = 0; r < n; r++) {
    const int kr = IN[r];
    for (int s = 0; s < n; s++) AS[r * P + s] = K[P * kr + IN[s]];
    AS[r * P + r] *= 1 + give[kr];
    RS[r] = wanted[kr];
  }
  for (int col = 0; col < n; col++) {
    const double pivot = AS[col * P + col];
    for (int r = col + 1; r < n; r++) {
      const double f = AS[r * P + col] / pivot;
      for (int s = col + 1; s < n; s++) AS[r * P + s] -= f * AS[col * P + s];
      RS[r] -= f * RS[col];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double x = RS[r];
    for (int s = r + 1; s < n; s++) x -= AS[r * P + s] * XS[s];
    XS[r] = x / AS[r * P + r];
    solved[IN[r]] = XS[r];
  }
}

// src/contact.js:1805-2090 (solveRows)
void solve_rows(OracleSoft* w, const Soft& soft, double h, bool bias, bool known, int from, int to) {
  double *v = w->velocity, *om = w->angular;
  const double* invI = w->inv_i;
  const double *p = w->position, *start = w->start, *q = w->rotation;
  const double inv = 1 / h;
  const double springy = soft.impulse_scale / soft.mass_scale + COMPLIANCE;
  double target[P], mass_scale[P], impulse_scale[P], give[P], closing_now[P], wanted[P], solved[P];
  double v3[3];
  for (int m = from; m < to; m++) {
    const int a = w->a[m], b = w->b[m], a3 = a * 3, b3 = b * 3, m3 = 3 * m;
    const double nx = w->normal[m3], ny = w->normal[m3 + 1], nz = w->normal[m3 + 2];
    const double dx = (p[b3] - start[b3]) - (p[a3] - start[a3]);
    const double dy = (p[b3 + 1] - start[b3 + 1]) - (p[a3 + 1] - start[a3 + 1]);
    const double dz = (p[b3 + 2] - start[b3 + 2]) - (p[a3 + 2] - start[a3 + 2]);
    const double max_bias = w->tolerance[m] * inv;
    const int count = w->points[m];
    double dn = 0, nax = 0, nay = 0, naz = 0, nbx = 0, nby = 0, nbz = 0;
    if (!known) {
      dn = dx * nx + dy * ny + dz * nz;
      rotate(q + a * 4, nx, ny, nz, true, v3);
      nax = v3[0]; nay = v3[1]; naz = v3[2];
      rotate(q + b * 4, nx, ny, nz, true, v3);
      nbx = v3[0]; nby = v3[1]; nbz = v3[2];
    }
    double least = INFINITY;
    for (int k = 0; k < count; k++) {
      const int i = m * P + k, i3 = 3 * i;
      double s;
      if (known) {
        s = w->now[i];
      } else {
        s = dn + (w->lb[i3] * nbx + w->lb[i3 + 1] * nby + w->lb[i3 + 2] * nbz)
            - (w->la[i3] * nax + w->la[i3 + 1] * nay + w->la[i3 + 2] * naz) + w->adjusted[i];
        if (!bias) w->now[i] = s;
      }
      if (s < least) least = s;
      if (s > 0) {
        target[k] = s * inv; mass_scale[k] = 1; impulse_scale[k] = 0; give[k] = COMPLIANCE;
      } else if (bias) {
        target[k] = js_max(soft.bias_rate * s, -max_bias);
        mass_scale[k] = soft.mass_scale; impulse_scale[k] = soft.impulse_scale; give[k] = springy;
      } else {
        target[k] = 0; mass_scale[k] = 1; impulse_scale[k] = 0; give[k] = COMPLIANCE;
      }
    }
    w->gap[m] = least;
    const double* jac = w->jac;
    const double ima = w->inv_mass[a], imb = w->inv_mass[b];
    const int kb = m * P * P;
    const double* K = w->block + kb;
    double vax = v[a3], vay = v[a3 + 1], vaz = v[a3 + 2], wax = om[a3], way = om[a3 + 1], waz = om[a3 + 2];
    double vbx = v[b3], vby = v[b3 + 1], vbz = v[b3 + 2], wbx = om[b3], wby = om[b3 + 1], wbz = om[b3 + 2];
    if (count == 1 || !(K[0] > 0)) {
      for (int k = 0; k < count; k++) {
        const int i = m * P + k, j = J * i;
        const double vn = (vbx - vax) * nx + (vby - vay) * ny + (vbz - vaz) * nz
            + wbx * jac[j + 3] + wby * jac[j + 4] + wbz * jac[j + 5] - wax * jac[j] - way * jac[j + 1]
            - waz * jac[j + 2];
        const double old = w->normal_impulse[i];
        double next = old - w->normal_mass[i] * mass_scale[k] * (vn + target[k]) - impulse_scale[k] * old;
        if (next < 0) next = 0;
        closing_now[k] = vn; solved[k] = next;
      }
    } else {
      for (int k = 0; k < count; k++) {
        const int j = J * (m * P + k);
        closing_now[k] = (vbx - vax) * nx + (vby - vay) * ny + (vbz - vaz) * nz
            + wbx * jac[j + 3] + wby * jac[j + 4] + wbz * jac[j + 5] - wax * jac[j] - way * jac[j + 1]
            - waz * jac[j + 2];
      }
      int set = 0;
      for (int k = 0; k < count; k++) {
        double r = -closing_now[k] - target[k];
        for (int l = 0; l < count; l++) r += K[P * k + l] * w->normal_impulse[m * P + l];
        wanted[k] = r;
        if (r > 0) set |= 1 << k;
      }
      for (int round = 0;; round++) {
        pushes(K, count, set, give, wanted, solved);
        int flip = -1;
        for (int k = 0; k < count && flip < 0; k++) {
          if ((set >> k) & 1) {
            if (solved[k] < 0) flip = k;
          } else {
            double r = -wanted[k];
            for (int l = 0; l < count; l++) r += K[P * k + l] * solved[l];
            if (r < 0) flip = k;
          }
        }
        if (flip < 0 || round == (1 << count)) break;
        set ^= 1 << flip;
      }
    }
    for (int k = 0; k < count; k++) {
      const int i = m * P + k, j = J * i;
      const double old = w->normal_impulse[i], next = solved[k] > 0 ? solved[k] : 0;
      const double d = next - old;
      w->normal_impulse[i] = next;
      if (d > 0 && w->max_normal_impulse[i] == 0) w->relative_velocity[i] = closing_now[k];
      if (d > w->max_normal_impulse[i]) w->max_normal_impulse[i] = d;
      const double px = nx * d, py = ny * d, pz = nz * d;
      vax -= ima * px; vay -= ima * py; vaz -= ima * pz;
      vbx += imb * px; vby += imb * py; vbz += imb * pz;
      wax -= jac[j + 6] * d; way -= jac[j + 7] * d; waz -= jac[j + 8] * d;
      wbx += jac[j + 9] * d; wby += jac[j + 10] * d; wbz += jac[j + 11] * d;
    }
    const double mu = w->friction[m];
    const int oa = a * 6, ob = b * 6;
    const double t1x = w->t1[m3], t1y = w->t1[m3 + 1], t1z = w->t1[m3 + 2];
    const double t2x = w->t2[m3], t2y = w->t2[m3 + 1], t2z = w->t2[m3 + 2];
    double pressed = 0, twisted = 0;
    for (int k = 0; k < count; k++) {
      const int i = m * P + k;
      pressed += w->normal_impulse[i];
      twisted += w->lever[i] * w->normal_impulse[i];
    }
    {
      const double* f = w->fjac + FJ * m;
      const double* g = f + 12;
      const double dvx = vbx - vax, dvy = vby - vay, dvz = vbz - vaz;
      double v1 = dvx * t1x + dvy * t1y + dvz * t1z
          + wbx * f[3] + wby * f[4] + wbz * f[5] - wax * f[0] - way * f[1] - waz * f[2];
      double v2 = dvx * t2x + dvy * t2y + dvz * t2z
          + wbx * g[3] + wby * g[4] + wbz * g[5] - wax * g[0] - way * g[1] - waz * g[2];
      const double s1 = w->drift1[m], s2 = w->drift2[m];
      if (bias && (s1 != 0 || s2 != 0)) {
        double b1 = soft.bias_rate * s1, b2 = soft.bias_rate * s2;
        const double bb = b1 * b1 + b2 * b2;
        if (bb > max_bias * max_bias) {
          const double scale = max_bias / std::sqrt(bb);
          b1 *= scale; b2 *= scale;
        }
        v1 += b1; v2 += b2;
      }
      const double* tm = w->tangent_mass + m3;
      const double o1 = w->tangent_impulse1[m], o2 = w->tangent_impulse2[m];
      double n1 = o1 - (tm[0] * v1 + tm[1] * v2), n2 = o2 - (tm[1] * v1 + tm[2] * v2);
      const double limit = mu * pressed, l2 = n1 * n1 + n2 * n2;
      if (l2 > limit * limit) {
        const double ma = tm[0], mb = tm[1], mc = tm[2];
        double al = 0, x = n1, y = n2, ll = l2;
        while (limit > 0) {
          const double ab = al * mb, d = (1 + al * ma) * (1 + al * mc) - ab * ab;
          const double px = ma * x + mb * y, py = mb * x + mc * y;
          const double wx = ((1 + al * mc) * px - ab * py) / d, wy = ((1 + al * ma) * py - ab * px) / d;
          const double lw = x * wx + y * wy;
          if (!(lw > 0)) break;
          const double next = al + (std::sqrt(ll) / limit - 1) * ll / lw;
          if (!(next > al)) break;
          al = next;
          const double ab2 = al * mb, d2 = (1 + al * ma) * (1 + al * mc) - ab2 * ab2;
          x = ((1 + al * mc) * n1 - ab2 * n2) / d2; y = ((1 + al * ma) * n2 - ab2 * n1) / d2;
          ll = x * x + y * y;
          if (!(ll > limit * limit)) break;
        }
        const double scale = limit / std::sqrt(ll);
        n1 = x * scale; n2 = y * scale;
        w->sliding[m] = 1;
      } else {
        w->sliding[m] = 0;
      }
      w->tangent_impulse1[m] = n1; w->tangent_impulse2[m] = n2;
      const double d1 = n1 - o1, d2 = n2 - o2;
      const double px = t1x * d1 + t2x * d2, py = t1y * d1 + t2y * d2, pz = t1z * d1 + t2z * d2;
      vax -= ima * px; vay -= ima * py; vaz -= ima * pz;
      vbx += imb * px; vby += imb * py; vbz += imb * pz;
      wax -= f[6] * d1 + g[6] * d2; way -= f[7] * d1 + g[7] * d2; waz -= f[8] * d1 + g[8] * d2;
      wbx += f[9] * d1 + g[9] * d2; wby += f[10] * d1 + g[10] * d2; wbz += f[11] * d1 + g[11] * d2;
    }
    if (count <= 2) {
      const double most = w->patch[m] * pressed;
      const double ux = wbx - wax, uy = wby - way, uz = wbz - waz;
      const double r1 = w->roll_impulse1[m], r2 = w->roll_impulse2[m];
      double n1 = r1 - w->roll_mass1[m] * (ux * t1x + uy * t1y + uz * t1z);
      double n2 = r2 - w->roll_mass2[m] * (ux * t2x + uy * t2y + uz * t2z);
      const double l2 = n1 * n1 + n2 * n2;
      if (l2 > most * most) {
        const double scale = most / std::sqrt(l2);
        n1 *= scale; n2 *= scale;
      }
      w->roll_impulse1[m] = n1; w->roll_impulse2[m] = n2;
      const double d1 = n1 - r1, d2 = n2 - r2;
      const double lx = t1x * d1 + t2x * d2, ly = t1y * d1 + t2y * d2, lz = t1z * d1 + t2z * d2;
      wax -= invI[oa] * lx + invI[oa + 1] * ly + invI[oa + 2] * lz;
      way -= invI[oa + 1] * lx + invI[oa + 3] * ly + invI[oa + 4] * lz;
      waz -= invI[oa + 2] * lx + invI[oa + 4] * ly + invI[oa + 5] * lz;
      wbx += invI[ob] * lx + invI[ob + 1] * ly + invI[ob + 2] * lz;
      wby += invI[ob + 1] * lx + invI[ob + 3] * ly + invI[ob + 4] * lz;
      wbz += invI[ob + 2] * lx + invI[ob + 4] * ly + invI[ob + 5] * lz;
    }
    {
      const double limit = count == 1 ? SPIN_DISC * mu * w->patch[m] * pressed : mu * twisted;
      const double s0 = w->spin_impulse[m];
      double s1 = s0 - w->spin_mass[m] * ((wbx - wax) * nx + (wby - way) * ny + (wbz - waz) * nz);
      if (s1 > limit) s1 = limit; else if (s1 < -limit) s1 = -limit;
      w->spin_impulse[m] = s1;
      const double ds = s1 - s0, lx = nx * ds, ly = ny * ds, lz = nz * ds;
      wax -= invI[oa] * lx + invI[oa + 1] * ly + invI[oa + 2] * lz;
      way -= invI[oa + 1] * lx + invI[oa + 3] * ly + invI[oa + 4] * lz;
      waz -= invI[oa + 2] * lx + invI[oa + 4] * ly + invI[oa + 5] * lz;
      wbx += invI[ob] * lx + invI[ob + 1] * ly + invI[ob + 2] * lz;
      wby += invI[ob + 1] * lx + invI[ob + 3] * ly + invI[ob + 4] * lz;
      wbz += invI[ob + 2] * lx + invI[ob + 4] * ly + invI[ob + 5] * lz;
    }
    write_back(v, om, ima, a3, vax, vay, vaz, wax, way, waz);
    write_back(v, om, imb, b3, vbx, vby, vbz, wbx, wby, wbz);
  }
}

// src/world.js:1495-1530 (_accelerate)
void accelerate(OracleSoft* w, double h) {
  const double gx = w->gravity[0] * h, gy = w->gravity[1] * h, gz = w->gravity[2] * h;
  const double most = SPIN_TURN / (h * w->substeps);
  double *v = w->velocity, *om = w->angular;
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (w->type[s] != DYNAMIC || w->frozen[s] != 0) continue;
    const int s3 = s * 3;
    const double g = w->gravity_scale[s];
    v[s3] += gx * g; v[s3 + 1] += gy * g; v[s3 + 2] += gz * g;
    if (w->damping[s] != 0) {
      const double k = 1 / (1 + h * w->damping[s]);
      v[s3] *= k; v[s3 + 1] *= k; v[s3 + 2] *= k;
    }
    if (w->angular_damping[s] != 0) {
      const double k = 1 / (1 + h * w->angular_damping[s]);
      om[s3] *= k; om[s3 + 1] *= k; om[s3 + 2] *= k;
    }
    if (w->shape[s] != SPHERE_SHAPE || w->framed[s] == 1) {
      const double w2 = om[s3] * om[s3] + om[s3 + 1] * om[s3 + 1] + om[s3 + 2] * om[s3 + 2];
      if (w2 > most * most) {
        const double k = most / std::sqrt(w2);
        om[s3] *= k; om[s3 + 1] *= k; om[s3 + 2] *= k;
      }
    }
  }
}

// src/world.js:2820-2828 (advance): out = normalize(q + h (w, 0) q).
void advance(const double* q, double wx, double wy, double wz, double h, double* out) {
  const double x = q[0], y = q[1], z = q[2], ww = q[3];
  const double nx = x + h * (wx * ww + wy * z - wz * y);
  const double ny = y + h * (wy * ww + wz * x - wx * z);
  const double nz = z + h * (wz * ww + wx * y - wy * x);
  const double nw = ww - h * (wx * x + wy * y + wz * z);
  const double l = 1 / std::sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
  out[0] = nx * l; out[1] = ny * l; out[2] = nz * l; out[3] = nw * l;
}

// src/world.js:2831-2834 (angular): v3 = R I^-1 R^T L.
void angular(const double* q, const double* inv, double lx, double ly, double lz, double* v3) {
  rotate(q, lx, ly, lz, true, v3);
  rotate(q, v3[0] * inv[0], v3[1] * inv[1], v3[2] * inv[2], false, v3);
}

// src/world.js:2777-2805 (untwist)
void untwist(double* q, const double* inv, double lx, double ly, double lz, double most) {
  double B[3];
  rotate(q, lx, ly, lz, true, B);
  const double bx = B[0], by = B[1], bz = B[2], ll = bx * bx + by * by + bz * bz;
  double x = bx, y = by, z = bz;
  for (int round = 0; round < 2; round++) {
    const double gx = x * inv[0], gy = y * inv[1], gz = z * inv[2];
    const double over = x * gx + y * gy + z * gz - most;
    if (!(over > 0)) break;
    const double k = (gx * x + gy * y + gz * z) / ll, tx = gx - k * x, ty = gy - k * y, tz = gz - k * z;
    const double tt = tx * tx + ty * ty + tz * tz;
    if (!(tt > 0)) break;
    const double a = 0.5 * over / tt;
    x -= a * tx; y -= a * ty; z -= a * tz;
    const double back = std::sqrt(ll / (x * x + y * y + z * z));
    x *= back; y *= back; z *= back;
  }
  const double cx = y * bz - z * by, cy = z * bx - x * bz, cz = x * by - y * bx;
  const double cw = ll + x * bx + y * by + z * bz;
  const double l = std::sqrt(cx * cx + cy * cy + cz * cz + cw * cw);
  if (!(l > 0)) return;
  const double T4[4] = {cx / l, cy / l, cz / l, cw / l};
  double Q4[4];
  quat_multiply(q, T4, Q4);
  const double n = 1 / std::sqrt(Q4[0] * Q4[0] + Q4[1] * Q4[1] + Q4[2] * Q4[2] + Q4[3] * Q4[3]);
  q[0] = Q4[0] * n; q[1] = Q4[1] * n; q[2] = Q4[2] * n; q[3] = Q4[3] * n;
}

// src/world.js:2739-2763 (tumble)
void tumble(OracleSoft* w, int s, double dt) {
  double* q = w->rotation + 4 * s;
  double* om = w->angular + 3 * s;
  const double* I = w->inertia + 3 * s;
  const double* inv = w->inv_inertia + 3 * s;
  double* left = w->tumbled + 3 * s;
  double v3[3], q4[4];
  rotate(q, om[0], om[1], om[2], true, v3);
  rotate(q, v3[0] * I[0], v3[1] * I[1], v3[2] * I[2], false, v3);
  const double lx = v3[0], ly = v3[1], lz = v3[2];
  if (om[0] != left[0] || om[1] != left[1] || om[2] != left[2]) {
    w->ceiling[s] = om[0] * lx + om[1] * ly + om[2] * lz;
  }
  const double most = w->ceiling[s];
  advance(q, om[0], om[1], om[2], dt * 0.25, q4);
  angular(q4, inv, lx, ly, lz, v3);
  advance(q, v3[0], v3[1], v3[2], dt * 0.5, q);
  angular(q, inv, lx, ly, lz, v3);
  if (v3[0] * lx + v3[1] * ly + v3[2] * lz > most * (1 + ROUNDED)) {
    untwist(q, inv, lx, ly, lz, most);
    angular(q, inv, lx, ly, lz, v3);
  }
  om[0] = v3[0]; om[1] = v3[1]; om[2] = v3[2];
  left[0] = v3[0]; left[1] = v3[1]; left[2] = v3[2];
}

// src/world.js:1533-1563 (_move)
void move(OracleSoft* w, double h) {
  const double half = h * 0.5;
  double *p = w->position, *v = w->velocity, *om = w->angular;
  const double* I = w->inertia;
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i], s3 = s * 3, s4 = s * 4;
    if (w->frozen[s] != 0) continue;
    p[s3] += v[s3] * h; p[s3 + 1] += v[s3 + 1] * h; p[s3 + 2] += v[s3 + 2] * h;
    const double wx = om[s3], wy = om[s3 + 1], wz = om[s3 + 2];
    const bool turning = wx != 0 || wy != 0 || wz != 0;
    if (turning && w->type[s] == DYNAMIC && w->lock[s] == 0 && w->joint_head[s] == 0
        && (I[s3] != I[s3 + 1] || I[s3 + 1] != I[s3 + 2])) {
      tumble(w, s, h);
    } else if (turning) {
      double* q = w->rotation + s4;
      advance(q, wx, wy, wz, half, q);
    }
  }
}

// src/contact.js:2216-2220 (closing) and 1743-1749 (push)
double closing(const double* v, const double* om, int a3, int b3, const double* jac, double dx, double dy, double dz) {
  return (v[b3] - v[a3]) * dx + (v[b3 + 1] - v[a3 + 1]) * dy + (v[b3 + 2] - v[a3 + 2]) * dz
      + om[b3] * jac[3] + om[b3 + 1] * jac[4] + om[b3 + 2] * jac[5]
      - om[a3] * jac[0] - om[a3 + 1] * jac[1] - om[a3 + 2] * jac[2];
}
void push(double* v, double* om, double ima, double imb, int a3, int b3, const double* jac, double dx, double dy,
          double dz, double l) {
  const double px = dx * l, py = dy * l, pz = dz * l;
  v[a3] -= ima * px; v[a3 + 1] -= ima * py; v[a3 + 2] -= ima * pz;
  v[b3] += imb * px; v[b3 + 1] += imb * py; v[b3 + 2] += imb * pz;
  om[a3] -= jac[6] * l; om[a3 + 1] -= jac[7] * l; om[a3 + 2] -= jac[8] * l;
  om[b3] += jac[9] * l; om[b3 + 1] += jac[10] * l; om[b3 + 2] += jac[11] * l;
}

// src/contact.js:2229-2251
void restitution(OracleSoft* w, double gravity) {
  double *v = w->velocity, *om = w->angular;
  for (int f = 0; f < w->count; f++) {
    const int m = w->row[f];
    const double e = w->restitution[m];
    if (e == 0 || m >= w->solved) continue;
    const double threshold = std::sqrt(2 * gravity * w->tolerance[m]);
    const int a = w->a[m], b = w->b[m], a3 = a * 3, b3 = b * 3, m3 = 3 * m;
    const double nx = w->normal[m3], ny = w->normal[m3 + 1], nz = w->normal[m3 + 2];
    for (int k = 0; k < w->points[m]; k++) {
      const int i = m * P + k;
      if (w->relative_velocity[i] > -threshold || w->max_normal_impulse[i] == 0) continue;
      const double vn = closing(v, om, a3, b3, w->jac + J * i, nx, ny, nz);
      const double old = w->normal_impulse[i];
      double next = old - w->normal_mass[i] * (vn + e * w->relative_velocity[i]);
      if (next < 0) next = 0;
      const double d = next - old;
      w->normal_impulse[i] = next;
      push(v, om, w->inv_mass[a], w->inv_mass[b], a3, b3, w->jac + J * i, nx, ny, nz, d);
    }
  }
}

// src/contact.js:1168-1182 (turned): the step's move and turn vector of slot s into T.
void turned(const OracleSoft* w, int s, double* T) {
  const double *p = w->position, *st = w->start, *q = w->rotation + 4 * s, *q0 = w->start_q + 4 * s;
  const int s3 = s * 3;
  const double ax = q[0], ay = q[1], az = q[2], aw = q[3];
  const double bx = -q0[0], by = -q0[1], bz = -q0[2], bw = q0[3];
  double ux = aw * bx + ax * bw + ay * bz - az * by;
  double uy = aw * by + ay * bw + az * bx - ax * bz;
  double uz = aw * bz + az * bw + ax * by - ay * bx;
  double uw = aw * bw - ax * bx - ay * by - az * bz;
  if (uw < 0) { ux = -ux; uy = -uy; uz = -uz; uw = -uw; }
  const double sine = std::sqrt(ux * ux + uy * uy + uz * uz);
  const double k = sine == 0 ? 0 : 2 * js_atan2(sine, uw) / sine;
  T[0] = p[s3] - st[s3]; T[1] = p[s3 + 1] - st[s3 + 1]; T[2] = p[s3 + 2] - st[s3 + 2];
  T[3] = ux * k; T[4] = uy * k; T[5] = uz * k;
}

// src/contact.js:1122-1153 (slipped)
void slipped(OracleSoft* w, int m) {
  const int a = w->a[m], b = w->b[m], m3 = 3 * m;
  double pressed = 0;
  bool impact = false;
  for (int k = 0; k < w->points[m]; k++) {
    const int i = m * P + k, i3 = 3 * i;
    pressed += w->normal_impulse[i];
    if (w->shift[i3] != 0 || w->shift[i3 + 1] != 0 || w->shift[i3 + 2] != 0) impact = true;
  }
  if (w->sliding[m] != 0 || pressed == 0 || impact) {
    w->drift1[m] = 0; w->drift2[m] = 0;
    return;
  }
  double T[12];
  double dx = 0, dy = 0, dz = 0;
  if (w->awake[b] != 0) {
    turned(w, b, T);
    const double rx = w->anchor_b[m3], ry = w->anchor_b[m3 + 1], rz = w->anchor_b[m3 + 2];
    dx = T[0] + T[4] * rz - T[5] * ry;
    dy = T[1] + T[5] * rx - T[3] * rz;
    dz = T[2] + T[3] * ry - T[4] * rx;
  }
  if (w->awake[a] != 0) {
    turned(w, a, T + 6);
    const double rx = w->anchor_a[m3], ry = w->anchor_a[m3 + 1], rz = w->anchor_a[m3 + 2];
    dx -= T[6] + T[10] * rz - T[11] * ry;
    dy -= T[7] + T[11] * rx - T[9] * rz;
    dz -= T[8] + T[9] * ry - T[10] * rx;
  }
  w->drift1[m] += dx * w->t1[m3] + dy * w->t1[m3 + 1] + dz * w->t1[m3 + 2];
  w->drift2[m] += dx * w->t2[m3] + dy * w->t2[m3 + 1] + dz * w->t2[m3 + 2];
}

// src/contact.js:1091-1115 (keep), into the pair list at p.
void keep(OracleSoft* w, int p, int m) {
  const int count = w->points[m], m3 = 3 * m, p3 = 3 * p;
  w->pair_points[p] = (uint8_t)count;
  const double t1x = w->t1[m3], t1y = w->t1[m3 + 1], t1z = w->t1[m3 + 2];
  const double t2x = w->t2[m3], t2y = w->t2[m3 + 1], t2z = w->t2[m3 + 2];
  const double r1 = w->roll_impulse1[m], r2 = w->roll_impulse2[m], sp = w->spin_impulse[m];
  w->pair_twist[p3] = t1x * r1 + t2x * r2 + w->normal[m3] * sp;
  w->pair_twist[p3 + 1] = t1y * r1 + t2y * r2 + w->normal[m3 + 1] * sp;
  w->pair_twist[p3 + 2] = t1z * r1 + t2z * r2 + w->normal[m3 + 2] * sp;
  const double l1 = w->tangent_impulse1[m], l2 = w->tangent_impulse2[m];
  w->pair_friction[p3] = t1x * l1 + t2x * l2;
  w->pair_friction[p3 + 1] = t1y * l1 + t2y * l2;
  w->pair_friction[p3 + 2] = t1z * l1 + t2z * l2;
  const double d1 = w->drift1[m], d2 = w->drift2[m];
  w->pair_slip[p3] = t1x * d1 + t2x * d2;
  w->pair_slip[p3 + 1] = t1y * d1 + t2y * d2;
  w->pair_slip[p3 + 2] = t1z * d1 + t2z * d2;
  for (int k = 0; k < count; k++) {
    const int i = m * P + k, o = p * P + k;
    w->pair_pid[o] = w->pid[i];
    w->pair_normal_impulse[o] = w->normal_impulse[i];
  }
}

// src/contact.js:1067-1088 (store): slip for every solved manifold; the pair list keeps the
// plain pairs' (level pairs' Memory is the caller's).
int store(OracleSoft* w) {
  for (int f = 0; f < w->count; f++) {
    const int m = w->row[f];
    if (m >= w->solved) continue;
    slipped(w, m);
    if (w->sub[m] >= 0) continue;   // a level pair's: into its Memory, which the caller keeps
    keep(w, w->pair[m], m);
  }
  return 0;
}

// The reference's workers (src/core/jobs.js; the prepare, warm and solve jobs of
// src/contact.js:1341, 1611, 1793, 2129-2146) as host threads: a parallel-for over a row
// range, every thread the calling one included claiming the next chunk from one atomic
// cursor. A dispatch of fewer rows than its job's least runs on the calling thread alone;
// otherwise the range is cut into (threads * 4) chunks. Each dispatch ends at a barrier of
// all threads, so no thread can carry one dispatch's arguments into the next. Between the
// step's dispatches the workers spin, as the reference's do from begin() to end().
enum Job { WARM = 0, SOLVE = 1, PREPARE = 2 };
constexpr int LEAST[3] = {16, 4, 2};   // src/contact.js:2146
constexpr int PER_THREAD = 4;          // src/core/jobs.js:94

inline void relax_cpu() {
#if defined(__x86_64__) || defined(__i386__)
  __builtin_ia32_pause();
#else
  std::this_thread::yield();
#endif
}

class Crew {
 public:
  Crew(OracleSoft* w, int workers) : w_(w), threads_(workers + 1) {
    for (int i = 0; i < workers; i++) pool_.emplace_back([this] { serve(); });
  }
  ~Crew() {
    quit_.store(true, std::memory_order_release);
    for (std::thread& t : pool_) t.join();
  }
  // Runs `job` over rows from..to-1, with the solve's arguments where it is a solve.
  void run(Job job, int from, int to, const Soft* soft = nullptr, double h = 0, bool bias = false,
           bool known = false) {
    const int n = to - from;
    if (n <= 0) return;
    job_ = job; from_ = from; to_ = to; soft_ = soft; h_ = h; bias_ = bias; known_ = known;
    if (pool_.empty() || n < LEAST[job]) {
      rows(from, to);
      return;
    }
    const int chunks = threads_ * PER_THREAD;
    chunk_ = (n + chunks - 1) / chunks;
    next_.store(0, std::memory_order_relaxed);
    arrived_.store(0, std::memory_order_relaxed);
    epoch_.fetch_add(1, std::memory_order_acq_rel);
    claim();
    arrived_.fetch_add(1, std::memory_order_acq_rel);
    while (arrived_.load(std::memory_order_acquire) < threads_) relax_cpu();
  }

 private:
  void rows(int from, int to) {
    switch (job_) {
      case WARM: warm_rows(w_, from, to); break;
      case SOLVE: solve_rows(w_, *soft_, h_, bias_, known_, from, to); break;
      case PREPARE: prepare(w_, from, to); break;
    }
  }
  void claim() {
    for (;;) {
      const int f = from_ + next_.fetch_add(1, std::memory_order_relaxed) * chunk_;
      if (f >= to_) return;
      rows(f, std::min(to_, f + chunk_));
    }
  }
  void serve() {
    int seen = 0;
    for (;;) {
      int e;
      while ((e = epoch_.load(std::memory_order_acquire)) == seen) {
        if (quit_.load(std::memory_order_acquire)) return;
        relax_cpu();
      }
      seen = e;
      claim();
      arrived_.fetch_add(1, std::memory_order_acq_rel);
    }
  }

  OracleSoft* w_;
  const int threads_;
  std::vector<std::thread> pool_;
  std::atomic<int> epoch_{0}, next_{0}, arrived_{0};
  std::atomic<bool> quit_{false};
  Job job_ = WARM;
  int from_ = 0, to_ = 0, chunk_ = 1;
  const Soft* soft_ = nullptr;
  double h_ = 0;
  bool bias_ = false, known_ = false;
};

// src/contact.js:1604-1613 (warmStart) and 1781-1796 (solve): colour by colour through the
// workers, then the overflow's rows, which may share bodies, on this thread in order. With
// no crew, every solved row in order on this thread (the rows lie colour by colour, so the
// order is the same).
void warm_pass(OracleSoft* w, Crew* crew) {
  if (crew == nullptr) { warm_rows(w, 0, w->solved); return; }
  for (int k = 0; k < COLOURS; k++) crew->run(WARM, w->starts[k], w->starts[k + 1]);
  warm_rows(w, w->starts[OVERFLOW], w->starts[OVERFLOW + 1]);
}
void solve_pass(OracleSoft* w, Crew* crew, const Soft& soft, double h, bool bias, bool known) {
  if (crew == nullptr) { solve_rows(w, soft, h, bias, known, 0, w->solved); return; }
  for (int k = 0; k < COLOURS; k++) crew->run(SOLVE, w->starts[k], w->starts[k + 1], &soft, h, bias, known);
  solve_rows(w, soft, h, bias, known, w->starts[OVERFLOW], w->starts[OVERFLOW + 1]);
}

// src/world.js:961-1031, for a world with no joints. Moving the bodies, the restitution pass
// and store stay on this thread, as the reference keeps them.
int step(OracleSoft* w, Crew* crew) {
  const bool touching = w->count > 0;
  const double h = w->dt / w->substeps;
  if (touching) inertias(w);
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    for (int k = 0; k < 3; k++) w->start[3 * s + k] = w->position[3 * s + k];
    for (int k = 0; k < 4; k++) w->start_q[4 * s + k] = w->rotation[4 * s + k];
  }
  Soft soft;
  if (touching) {
    if (crew != nullptr) crew->run(PREPARE, 0, w->solved);
    else prepare(w, 0, w->solved);
    soft = make_soft(STIFFNESS / h, DAMPING_RATIO, h);
  }
  for (int sub = 0; sub < w->substeps; sub++) {
    accelerate(w, h);
    if (touching) warm_pass(w, crew);
    if (touching) solve_pass(w, crew, soft, h, true, sub > 0);
    move(w, h);
    if (touching) solve_pass(w, crew, soft, h, false, false);
  }
  if (touching) {
    const double* g = w->gravity;
    restitution(w, std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
    return store(w);
  }
  (void)RESTING;
  return 0;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" int oracle_soft_step(OracleSoft* w) { return step(w, nullptr); }

extern "C" int oracle_soft_step_workers(OracleSoft* w, int workers) {
  if (workers <= 0) return step(w, nullptr);
  Crew crew(w, workers);
  return step(w, &crew);
}
