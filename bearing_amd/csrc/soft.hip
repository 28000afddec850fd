// The soft step on an MI355X: one workgroup per world, persistent across the
// whole phase (inertias, prepare, every substep's integrate / warm start / solve /
// relax, restitution, store), with a workgroup barrier between colours.
//
// Why this shape. A step's solve is 8 substeps x 3 sweeps x (colours + overflow)
// dependent sweeps (about 216 for nine colours), each a few hundred manifolds for a
// small world: a kernel launch per sweep would cost more than the sweep. Inside one
// workgroup a sweep ends at a barrier instead. A batch of worlds is a grid of
// workgroups, one world each, which fills the 256 CUs when there are thousands of
// worlds. Each thread takes whole manifolds (up to four points, the reference's
// block LCP is per manifold), gathering its two bodies' velocities from HBM and
// scattering them back; within a colour no two manifolds share a body the solve
// moves (src/graph.js:1-16), so the scatters never race, and static, frozen and
// kinematic bodies (inverse mass 0) are never written (src/contact.js:2096-2104).
//
// Arithmetic is the reference's, operation for operation, in doubles, built with
// -ffp-contract=off: the results are the reference's bits. Each function cites
// the lines it follows.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jsmath.h"
#include "soft.h"

namespace bm {

constexpr int P = 4;      // points per manifold, src/contact.js:128
constexpr int J = 12;     // turning terms per point, src/contact.js:139
constexpr int FJ = 24;
constexpr int DYNAMIC = 0;
constexpr int SPHERE_SHAPE = 1, CAPSULE_SHAPE = 2;
constexpr int COLOURS = 32, OVERFLOW = COLOURS;
constexpr double STIFFNESS = 1.0 / 8;               // src/world.js:115
constexpr double DAMPING_RATIO = 10;                // src/world.js:118
constexpr double SPIN_TURN = 3.141592653589793;     // src/world.js:171
constexpr double SPIN_DISC = 2.0 / 3;               // src/contact.js:146
constexpr double COMPLIANCE = 1.0 / 1048576;        // 2^-20, src/forest.js:67
constexpr double ROUNDED = 1.0 / 17592186044416.0;  // 2^-44, src/world.js:2809
constexpr double MATH_PI = 3.141592653589793;

struct Soft {
  double bias_rate, mass_scale, impulse_scale;
};

// src/contact.js:1590-1597
__device__ __forceinline__ Soft make_soft(double hertz, double zeta, double h) {
  const double omega = 2 * MATH_PI * hertz;
  const double a1 = 2 * zeta + h * omega, a2 = h * omega * a1, a3 = 1 / (1 + a2);
  return Soft{omega / a1, a2 * a3, a3};
}

// src/contact.js:1520-1549 (lockedInverse)
__device__ __forceinline__ void locked_inverse(const double* ax, const double* I, int lock, double* out) {
  const double i0 = I[0], i1 = I[1], i2 = I[2];
  const double x0 = ax[0], x1 = ax[1], x2 = ax[2];
  const double y0 = ax[3], y1 = ax[4], y2 = ax[5];
  const double z0 = ax[6], z1 = ax[7], z2 = ax[8];
  double M[6];
  M[0] = x0 * x0 * i0 + y0 * y0 * i1 + z0 * z0 * i2;
  M[1] = x0 * x1 * i0 + y0 * y1 * i1 + z0 * z1 * i2;
  M[2] = x0 * x2 * i0 + y0 * y2 * i1 + z0 * z2 * i2;
  M[3] = x1 * x1 * i0 + y1 * y1 * i1 + z1 * z1 * i2;
  M[4] = x1 * x2 * i0 + y1 * y2 * i1 + z1 * z2 * i2;
  M[5] = x2 * x2 * i0 + y2 * y2 * i1 + z2 * z2 * i2;
  for (int k = 0; k < 6; k++) out[k] = 0;
  int first = -1, second = -1;
  for (int k = 0; k < 3; k++) {
    if ((lock >> k) & 1) continue;
    if (first < 0) first = k; else second = k;
  }
  if (first < 0) return;
  // Entry (r, c) of a symmetric 3x3 in xx xy xz yy yz zz: diagonal 0, 3, 5.
  const int ff = first == 0 ? 0 : first == 1 ? 3 : 5;
  if (second < 0) {
    if (M[ff] > 0) out[ff] = 1 / M[ff];
    return;
  }
  const int ss = second == 1 ? 3 : 5;
  const int fs = first == 0 ? (second == 1 ? 1 : 2) : 4;
  const double det = M[ff] * M[ss] - M[fs] * M[fs];
  if (!(det > 0)) return;
  out[ff] = M[ss] / det; out[ss] = M[ff] / det; out[fs] = -M[fs] / det;
}

// src/contact.js:1302-1310 and 1493-1508: R diag(1/I) R^T for an awake dynamic body.
__device__ __forceinline__ void inertia_of(const SoftDev& d, int g) {
  double* out = d.inv_i + 6 * g;
  if (d.frozen[g] != 0 || d.type[g] != DYNAMIC) {
    for (int k = 0; k < 6; k++) out[k] = 0;
    return;
  }
  const double* ax = d.axes + 9 * g;
  if (d.lock[g] != 0) {
    double o[6];
    locked_inverse(ax, d.inertia + 3 * g, d.lock[g], o);
    for (int k = 0; k < 6; k++) out[k] = o[k];
    return;
  }
  const double* inv = d.inv_inertia + 3 * g;
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

// src/contact.js:1709-1722 (turning, along)
__device__ __forceinline__ double along(const double* I, double x, double y, double z) {
  return I[0] * x * x + I[3] * y * y + I[5] * z * z + 2 * (I[1] * x * y + I[2] * x * z + I[4] * y * z);
}
__device__ __forceinline__ double turning(const double* Ia, const double* Ib, double x, double y, double z) {
  const double k = along(Ia, x, y, z) + along(Ib, x, y, z);
  return k > 0 ? 1 / k : 0;
}

// src/contact.js:1565-1581 (lever): the 12 turning terms into jac, the effective mass back.
__device__ __forceinline__ double lever(double* jac, const double* Ia, const double* Ib, double ima, double imb, double rax,
                               double ray, double raz, double rbx, double rby, double rbz, double dx, double dy,
                               double dz) {
  const double cax = ray * dz - raz * dy, cay = raz * dx - rax * dz, caz = rax * dy - ray * dx;
  const double cbx = rby * dz - rbz * dy, cby = rbz * dx - rbx * dz, cbz = rbx * dy - rby * dx;
  const double iax = Ia[0] * cax + Ia[1] * cay + Ia[2] * caz;
  const double iay = Ia[1] * cax + Ia[3] * cay + Ia[4] * caz;
  const double iaz = Ia[2] * cax + Ia[4] * cay + Ia[5] * caz;
  const double ibx = Ib[0] * cbx + Ib[1] * cby + Ib[2] * cbz;
  const double iby = Ib[1] * cbx + Ib[3] * cby + Ib[4] * cbz;
  const double ibz = Ib[2] * cbx + Ib[4] * cby + Ib[5] * cbz;
  jac[0] = cax; jac[1] = cay; jac[2] = caz;
  jac[3] = cbx; jac[4] = cby; jac[5] = cbz;
  jac[6] = iax; jac[7] = iay; jac[8] = iaz;
  jac[9] = ibx; jac[10] = iby; jac[11] = ibz;
  const double k = ima + imb + cax * iax + cay * iay + caz * iaz + cbx * ibx + cby * iby + cbz * ibz;
  return k > 0 ? 1 / k : 0;
}

// src/contact.js:1199-1220 (core) for a whole body: anchor in its frame into out;
// returns the anchor's world offset at the start.
__device__ __forceinline__ V3 core(const SoftDev& d, int g, double x, double y, double z, double* out) {
  const int shape = d.shape[g];
  if (shape == SPHERE_SHAPE) {
    out[0] = 0; out[1] = 0; out[2] = 0;
    return V3{0, 0, 0};
  }
  const double* q = d.rotation + 4 * g;
  V3 v = rotate(q, x, y, z, true);
  if (shape != CAPSULE_SHAPE) {
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
    return V3{x, y, z};
  }
  const double half = (d.dims[3 * g + 1] - 2 * d.dims[3 * g]) * 0.5;
  const double c = v.y < -half ? -half : v.y > half ? half : v.y;
  out[0] = 0; out[1] = c; out[2] = 0;
  return rotate(q, 0, c, 0, false);
}

// Where a world's body velocities live during the soft step: in LDS (v then w, six
// doubles a body, by the world's own slot) when the world fits, else in the
// global columns. Every per-body velocity access goes through this.
struct VW {
  double* vb;
  double* wb;
  int stride;
  __device__ __forceinline__ double* v(int slot) const { return vb + stride * slot; }
  __device__ __forceinline__ double* w(int slot) const { return wb + stride * slot; }
};

// src/contact.js:1346-1445 (prepareRows) for manifold row m (global), bodies ga, gb.
__device__ __forceinline__ void prepare_row(const SoftDev& d, const VW& vw, int m, int ga, int gb, int sa, int sb) {
  const double* Ia = d.inv_i + 6 * ga;
  const double* Ib = d.inv_i + 6 * gb;
  const double *pa = d.position + 3 * ga, *pb = d.position + 3 * gb;
  const double *va = vw.v(sa), *vb = vw.v(sb);
  const double *wa = vw.w(sa), *wb = vw.w(sb);
  const double ima = d.inv_mass[ga], imb = d.inv_mass[gb];
  const int m3 = 3 * m;
  const double nx = d.normal[m3], ny = d.normal[m3 + 1], nz = d.normal[m3 + 2];
  const int count = d.points[m];
  if (count <= 2) {
    d.roll_mass1[m] = turning(Ia, Ib, d.t1[m3], d.t1[m3 + 1], d.t1[m3 + 2]);
    d.roll_mass2[m] = turning(Ia, Ib, d.t2[m3], d.t2[m3 + 1], d.t2[m3 + 2]);
  }
  d.spin_mass[m] = turning(Ia, Ib, nx, ny, nz);
  const double tol = d.tolerance[m];
  double weight = 0, ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  double ex = 0, ey = 0, ez = 0, fx = 0, fy = 0, fz = 0;
  for (int k = 0; k < count; k++) {
    const int i = m * P + k, i3 = 3 * i;
    const double px = d.point[i3], py = d.point[i3 + 1], pz = d.point[i3 + 2];
    const double rax = px - pa[0], ray = py - pa[1], raz = pz - pa[2];
    const double rbx = px + d.shift[i3] - pb[0], rby = py + d.shift[i3 + 1] - pb[1], rbz = pz + d.shift[i3 + 2] - pb[2];
    d.ra[i3] = rax; d.ra[i3 + 1] = ray; d.ra[i3 + 2] = raz;
    d.rb[i3] = rbx; d.rb[i3 + 1] = rby; d.rb[i3 + 2] = rbz;
    const V3 ca = core(d, ga, rax, ray, raz, d.la + i3);
    const V3 cb = core(d, gb, rbx, rby, rbz, d.lb + i3);
    d.adjusted[i] = d.separation[i] - ((cb.x - ca.x) * nx + (cb.y - ca.y) * ny + (cb.z - ca.z) * nz);
    d.normal_mass[i] = lever(d.jac + J * i, Ia, Ib, ima, imb, rax, ray, raz, rbx, rby, rbz, nx, ny, nz);
    d.max_normal_impulse[i] = 0;
    const double s = d.separation[i] / tol, near = s <= 1 ? 1 : s >= 2 ? 0 : 2 - s;
    weight += near;
    ax += near * rax; ay += near * ray; az += near * raz; bx += near * rbx; by += near * rby; bz += near * rbz;
    ex += rax; ey += ray; ez += raz; fx += rbx; fy += rby; fz += rbz;
    const double dvx = vb[0] + wb[1] * rbz - wb[2] * rby - va[0] - wa[1] * raz + wa[2] * ray;
    const double dvy = vb[1] + wb[2] * rbx - wb[0] * rbz - va[1] - wa[2] * rax + wa[0] * raz;
    const double dvz = vb[2] + wb[0] * rby - wb[1] * rbx - va[2] - wa[0] * ray + wa[1] * rax;
    d.relative_velocity[i] = dvx * nx + dvy * ny + dvz * nz;
  }
  if (count >= 2) {
    const double* jac = d.jac + J * (m * P);
    double* K = d.block + m * P * P;
    const double s = ima + imb;
    for (int k = 0; k < count; k++) {
      const double* jk = jac + J * k;
      for (int l = 0; l <= k; l++) {
        const double* jl = jac + J * l;
        const double e = s + jk[0] * jl[6] + jk[1] * jl[7] + jk[2] * jl[8] + jk[3] * jl[9] + jk[4] * jl[10]
            + jk[5] * jl[11];
        K[P * k + l] = e; K[P * l + k] = e;
      }
    }
  }
  if (weight == 0) { weight = count; ax = ex; ay = ey; az = ez; bx = fx; by = fy; bz = fz; }
  ax /= weight; ay /= weight; az /= weight; bx /= weight; by /= weight; bz /= weight;
  d.anchor_a[m3] = ax; d.anchor_a[m3 + 1] = ay; d.anchor_a[m3 + 2] = az;
  d.anchor_b[m3] = bx; d.anchor_b[m3 + 1] = by; d.anchor_b[m3 + 2] = bz;
  for (int k = 0; k < count; k++) {
    const int i3 = 3 * (m * P + k);
    double dx = d.ra[i3] - ax, dy = d.ra[i3 + 1] - ay, dz = d.ra[i3 + 2] - az;
    const double al = dx * nx + dy * ny + dz * nz;
    dx -= al * nx; dy -= al * ny; dz -= al * nz;
    d.lever[m * P + k] = dsqrt(dx * dx + dy * dy + dz * dz);
  }
  double* f = d.fjac + FJ * m;
  lever(f, Ia, Ib, ima, imb, ax, ay, az, bx, by, bz, d.t1[m3], d.t1[m3 + 1], d.t1[m3 + 2]);
  lever(f + 12, Ia, Ib, ima, imb, ax, ay, az, bx, by, bz, d.t2[m3], d.t2[m3 + 1], d.t2[m3 + 2]);
  const double k11 = ima + imb + f[0] * f[6] + f[1] * f[7] + f[2] * f[8] + f[3] * f[9] + f[4] * f[10] + f[5] * f[11];
  const double k22 = ima + imb + f[12] * f[18] + f[13] * f[19] + f[14] * f[20] + f[15] * f[21] + f[16] * f[22]
      + f[17] * f[23];
  const double k12 = f[12] * f[6] + f[13] * f[7] + f[14] * f[8] + f[15] * f[9] + f[16] * f[10] + f[17] * f[11];
  const double det = k11 * k22 - k12 * k12;
  double* tm = d.tangent_mass + m3;
  if (det > 0) {
    tm[0] = k22 / det; tm[1] = -k12 / det; tm[2] = k11 / det;
  } else {
    tm[0] = 0; tm[1] = 0; tm[2] = 0;
  }
}

// The two bodies' velocities, held in registers for a manifold (src/contact.js:1897-1908).
struct Pair {
  double vax, vay, vaz, wax, way, waz, vbx, vby, vbz, wbx, wby, wbz;
};

__device__ __forceinline__ Pair gather(const VW& vw, int sa, int sb) {
  const double *va = vw.v(sa), *wa = vw.w(sa), *vb = vw.v(sb), *wb = vw.w(sb);
  return Pair{va[0], va[1], va[2], wa[0], wa[1], wa[2], vb[0], vb[1], vb[2], wb[0], wb[1], wb[2]};
}

// src/contact.js:2096-2104 (writeBack): a body that cannot move is never written.
__device__ __forceinline__ void scatter(const VW& vw, int sa, int sb, double ima, double imb, const Pair& s) {
  if (ima != 0) {
    double *v = vw.v(sa), *w = vw.w(sa);
    v[0] = s.vax; v[1] = s.vay; v[2] = s.vaz; w[0] = s.wax; w[1] = s.way; w[2] = s.waz;
  }
  if (imb != 0) {
    double *v = vw.v(sb), *w = vw.w(sb);
    v[0] = s.vbx; v[1] = s.vby; v[2] = s.vbz; w[0] = s.wbx; w[1] = s.wby; w[2] = s.wbz;
  }
}

// src/contact.js:1616-1663 (warmRows) for row m of N points.
template <int N>
__device__ __forceinline__ void warm_row(const SoftDev& d, const VW& vw, int m, int ga, int gb, int sa, int sb) {
  const double ima = d.inv_mass[ga], imb = d.inv_mass[gb];
  const int m3 = 3 * m;
  const double nx = d.normal[m3], ny = d.normal[m3 + 1], nz = d.normal[m3 + 2];
  const double t1x = d.t1[m3], t1y = d.t1[m3 + 1], t1z = d.t1[m3 + 2];
  const double t2x = d.t2[m3], t2y = d.t2[m3 + 1], t2z = d.t2[m3 + 2];
  Pair s = gather(vw, sa, sb);
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int i = m * P + k;
    const double* jac = d.jac + J * i;
    const double ln = d.normal_impulse[i];
    const double px = nx * ln, py = ny * ln, pz = nz * ln;
    s.vax -= ima * px; s.vay -= ima * py; s.vaz -= ima * pz;
    s.vbx += imb * px; s.vby += imb * py; s.vbz += imb * pz;
    s.wax -= jac[6] * ln; s.way -= jac[7] * ln; s.waz -= jac[8] * ln;
    s.wbx += jac[9] * ln; s.wby += jac[10] * ln; s.wbz += jac[11] * ln;
  }
  {
    const double* fj = d.fjac + FJ * m;
    const double l1 = d.tangent_impulse1[m], l2 = d.tangent_impulse2[m];
    const double px = t1x * l1 + t2x * l2, py = t1y * l1 + t2y * l2, pz = t1z * l1 + t2z * l2;
    s.vax -= ima * px; s.vay -= ima * py; s.vaz -= ima * pz;
    s.vbx += imb * px; s.vby += imb * py; s.vbz += imb * pz;
    s.wax -= fj[6] * l1 + fj[18] * l2; s.way -= fj[7] * l1 + fj[19] * l2; s.waz -= fj[8] * l1 + fj[20] * l2;
    s.wbx += fj[9] * l1 + fj[21] * l2; s.wby += fj[10] * l1 + fj[22] * l2; s.wbz += fj[11] * l1 + fj[23] * l2;
  }
  {
    const double r1 = d.roll_impulse1[m], r2 = d.roll_impulse2[m], sp = d.spin_impulse[m];
    const double *Ia = d.inv_i + 6 * ga, *Ib = d.inv_i + 6 * gb;
    const double lx = t1x * r1 + t2x * r2 + nx * sp, ly = t1y * r1 + t2y * r2 + ny * sp, lz = t1z * r1 + t2z * r2 + nz * sp;
    s.wax -= Ia[0] * lx + Ia[1] * ly + Ia[2] * lz;
    s.way -= Ia[1] * lx + Ia[3] * ly + Ia[4] * lz;
    s.waz -= Ia[2] * lx + Ia[4] * ly + Ia[5] * lz;
    s.wbx += Ib[0] * lx + Ib[1] * ly + Ib[2] * lz;
    s.wby += Ib[1] * lx + Ib[3] * ly + Ib[4] * lz;
    s.wbz += Ib[2] * lx + Ib[4] * ly + Ib[5] * lz;
  }
  scatter(vw, sa, sb, ima, imb, s);
}

// src/contact.js:1672-1724 (pushes) for the points in `set` (bit k for point k), the
// rest 0: Gaussian elimination without pivoting on the set's rows and columns in
// rising order, exactly the reference's compacted loops (and its unrolled four-point
// case, the same operations), with every index known at compile time so the matrix
// stays in registers.
template <int N>
__device__ __forceinline__ void pushes(const double (&K)[N][N], int set, const double (&give)[N],
                                       const double (&wanted)[N], double (&solved)[N]) {
  double A[N][N], R[N], X[N];
#pragma unroll
  for (int r = 0; r < N; r++) {
#pragma unroll
    for (int c = 0; c < N; c++) A[r][c] = K[r][c];
    A[r][r] *= 1 + give[r];
    R[r] = wanted[r];
    solved[r] = 0;
  }
#pragma unroll
  for (int col = 0; col < N; col++) {
    if (!((set >> col) & 1)) continue;
    const double pivot = A[col][col];
#pragma unroll
    for (int r = col + 1; r < N; r++) {
      if (!((set >> r) & 1)) continue;
      const double f = A[r][col] / pivot;
#pragma unroll
      for (int c = col + 1; c < N; c++) {
        if ((set >> c) & 1) A[r][c] -= f * A[col][c];
      }
      R[r] -= f * R[col];
    }
  }
#pragma unroll
  for (int r = N - 1; r >= 0; r--) {
    if (!((set >> r) & 1)) continue;
    double x = R[r];
#pragma unroll
    for (int c = r + 1; c < N; c++) {
      if ((set >> c) & 1) x -= A[r][c] * X[c];
    }
    X[r] = x / A[r][r];
    solved[r] = X[r];
  }
}

// src/contact.js:1805-2090 (solveRows) for row m of N points.
template <int N>
__device__ __forceinline__ void solve_row(const SoftDev& d, const VW& vw, const Soft& soft, double h, bool bias, bool known, int m,
                          int ga, int gb, int sa, int sb) {
  const double inv = 1 / h;
  const double springy = soft.impulse_scale / soft.mass_scale + COMPLIANCE;
  double target[N], mass_scale[N], impulse_scale[N], give[N], closing_now[N], wanted[N], solved[N], impulse[N];
  double now[N], pushed[N];
  uint32_t first = 0, larger = 0;
  // Every input of the row is read before anything is written: the columns are
  // distinct arrays, and with no store in between the loads can all be in flight
  // together instead of one dependent round trip after another.
  const double tol = d.tolerance[m], mu = d.friction[m], patch = d.patch[m];
  const double s1_in = d.drift1[m], s2_in = d.drift2[m];
  const double o1 = d.tangent_impulse1[m], o2 = d.tangent_impulse2[m];
  const double r1_in = N <= 2 ? d.roll_impulse1[m] : 0, r2_in = N <= 2 ? d.roll_impulse2[m] : 0;
  const double rm1 = N <= 2 ? d.roll_mass1[m] : 0, rm2 = N <= 2 ? d.roll_mass2[m] : 0;
  const double s0_in = d.spin_impulse[m], spin_mass = d.spin_mass[m];
  const double tm0 = d.tangent_mass[3 * m], tm1 = d.tangent_mass[3 * m + 1], tm2 = d.tangent_mass[3 * m + 2];
  double most_in[N], lever_in[N], nmass[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    most_in[k] = d.max_normal_impulse[m * P + k];
    lever_in[k] = d.lever[m * P + k];
    nmass[k] = d.normal_mass[m * P + k];
  }
  const int m3 = 3 * m;
  const double nx = d.normal[m3], ny = d.normal[m3 + 1], nz = d.normal[m3 + 2];
  const double max_bias = tol * inv;
  double dn = 0;
  V3 na{0, 0, 0}, nb{0, 0, 0};
  if (!known) {
    const double *pa = d.position + 3 * ga, *pb = d.position + 3 * gb;
    const double *qa = d.start + 3 * ga, *qb = d.start + 3 * gb;
    const double dx = (pb[0] - qb[0]) - (pa[0] - qa[0]);
    const double dy = (pb[1] - qb[1]) - (pa[1] - qa[1]);
    const double dz = (pb[2] - qb[2]) - (pa[2] - qa[2]);
    dn = dx * nx + dy * ny + dz * nz;
    na = rotate(d.rotation + 4 * ga, nx, ny, nz, true);
    nb = rotate(d.rotation + 4 * gb, nx, ny, nz, true);
  }
  double least = __builtin_inf();
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int i = m * P + k, i3 = 3 * i;
    double s;
    if (known) {
      s = d.now[i];
    } else {
      s = dn + (d.lb[i3] * nb.x + d.lb[i3 + 1] * nb.y + d.lb[i3 + 2] * nb.z)
          - (d.la[i3] * na.x + d.la[i3 + 1] * na.y + d.la[i3 + 2] * na.z) + d.adjusted[i];
    }
    now[k] = s;
    if (s < least) least = s;
    if (s > 0) {
      target[k] = s * inv; mass_scale[k] = 1; impulse_scale[k] = 0; give[k] = COMPLIANCE;
    } else if (bias) {
      target[k] = js_max(soft.bias_rate * s, -max_bias);
      mass_scale[k] = soft.mass_scale; impulse_scale[k] = soft.impulse_scale; give[k] = springy;
    } else {
      target[k] = 0; mass_scale[k] = 1; impulse_scale[k] = 0; give[k] = COMPLIANCE;
    }
    impulse[k] = d.normal_impulse[i];
  }
  const double ima = d.inv_mass[ga], imb = d.inv_mass[gb];
  Pair s = gather(vw, sa, sb);
  const double* Kg = d.block + m * P * P;
  const bool single = N == 1 || !(Kg[0] > 0);
  if (single) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      const int i = m * P + k;
      const double* jac = d.jac + J * i;
      const double vn = (s.vbx - s.vax) * nx + (s.vby - s.vay) * ny + (s.vbz - s.vaz) * nz
          + s.wbx * jac[3] + s.wby * jac[4] + s.wbz * jac[5] - s.wax * jac[0] - s.way * jac[1] - s.waz * jac[2];
      const double old = impulse[k];
      double next = old - nmass[k] * mass_scale[k] * (vn + target[k]) - impulse_scale[k] * old;
      if (next < 0) next = 0;
      closing_now[k] = vn; solved[k] = next;
    }
  } else {
    double K[N][N];
#pragma unroll
    for (int k = 0; k < N; k++) {
#pragma unroll
      for (int l = 0; l < N; l++) K[k][l] = Kg[P * k + l];
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      const double* jac = d.jac + J * (m * P + k);
      closing_now[k] = (s.vbx - s.vax) * nx + (s.vby - s.vay) * ny + (s.vbz - s.vaz) * nz
          + s.wbx * jac[3] + s.wby * jac[4] + s.wbz * jac[5] - s.wax * jac[0] - s.way * jac[1] - s.waz * jac[2];
    }
    int set = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      double r = -closing_now[k] - target[k];
#pragma unroll
      for (int l = 0; l < N; l++) r += K[k][l] * impulse[l];
      wanted[k] = r;
      if (r > 0) set |= 1 << k;
    }
    // Murty's least-index pivoting, at most 2^N rounds (src/contact.js:1935-1951).
    for (int round = 0;; round++) {
      pushes<N>(K, set, give, wanted, solved);
      int flip = -1;
#pragma unroll
      for (int k = 0; k < N; k++) {
        if (flip >= 0) continue;
        if ((set >> k) & 1) {
          if (solved[k] < 0) flip = k;
        } else {
          double r = -wanted[k];
#pragma unroll
          for (int l = 0; l < N; l++) r += K[k][l] * solved[l];
          if (r < 0) flip = k;
        }
      }
      if (flip < 0 || round == (1 << N)) break;
      set ^= 1 << flip;
    }
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int i = m * P + k;
    const double* jac = d.jac + J * i;
    const double old = impulse[k], next = solved[k] > 0 ? solved[k] : 0;
    const double dd = next - old;
    impulse[k] = next;
    // The first push this step records the closing speed; the largest push is kept.
    if (dd > 0 && most_in[k] == 0) first |= 1u << k;
    if (dd > most_in[k]) { larger |= 1u << k; pushed[k] = dd; }
    const double px = nx * dd, py = ny * dd, pz = nz * dd;
    s.vax -= ima * px; s.vay -= ima * py; s.vaz -= ima * pz;
    s.vbx += imb * px; s.vby += imb * py; s.vbz += imb * pz;
    s.wax -= jac[6] * dd; s.way -= jac[7] * dd; s.waz -= jac[8] * dd;
    s.wbx += jac[9] * dd; s.wby += jac[10] * dd; s.wbz += jac[11] * dd;
  }
  const double *Ia = d.inv_i + 6 * ga, *Ib = d.inv_i + 6 * gb;
  const double t1x = d.t1[m3], t1y = d.t1[m3 + 1], t1z = d.t1[m3 + 2];
  const double t2x = d.t2[m3], t2y = d.t2[m3 + 1], t2z = d.t2[m3 + 2];
  double pressed = 0, twisted = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    pressed += impulse[k];
    twisted += lever_in[k] * impulse[k];
  }
  uint8_t sliding = 0;
  double t1_out, t2_out, r1_out = 0, r2_out = 0, s_out;
  {
    const double* f = d.fjac + FJ * m;
    const double* g = f + 12;
    const double dvx = s.vbx - s.vax, dvy = s.vby - s.vay, dvz = s.vbz - s.vaz;
    double v1 = dvx * t1x + dvy * t1y + dvz * t1z
        + s.wbx * f[3] + s.wby * f[4] + s.wbz * f[5] - s.wax * f[0] - s.way * f[1] - s.waz * f[2];
    double v2 = dvx * t2x + dvy * t2y + dvz * t2z
        + s.wbx * g[3] + s.wby * g[4] + s.wbz * g[5] - s.wax * g[0] - s.way * g[1] - s.waz * g[2];
    const double s1 = s1_in, s2 = s2_in;
    if (bias && (s1 != 0 || s2 != 0)) {
      double b1 = soft.bias_rate * s1, b2 = soft.bias_rate * s2;
      const double bb = b1 * b1 + b2 * b2;
      if (bb > max_bias * max_bias) {
        const double scale = max_bias / dsqrt(bb);
        b1 *= scale; b2 *= scale;
      }
      v1 += b1; v2 += b2;
    }
    double n1 = o1 - (tm0 * v1 + tm1 * v2), n2 = o2 - (tm1 * v1 + tm2 * v2);
    const double limit = mu * pressed, l2 = n1 * n1 + n2 * n2;
    if (l2 > limit * limit) {
      // Newton's method on the trust-region multiplier, as the reference does (src/contact.js:1987-2030).
      const double ma = tm0, mb = tm1, mc = tm2;
      double al = 0, x = n1, y = n2, ll = l2;
      while (limit > 0) {
        const double ab = al * mb, det = (1 + al * ma) * (1 + al * mc) - ab * ab;
        const double px = ma * x + mb * y, py = mb * x + mc * y;
        const double wx = ((1 + al * mc) * px - ab * py) / det, wy = ((1 + al * ma) * py - ab * px) / det;
        const double lw = x * wx + y * wy;
        if (!(lw > 0)) break;
        const double next = al + (dsqrt(ll) / limit - 1) * ll / lw;
        if (!(next > al)) break;
        al = next;
        const double ab2 = al * mb, d2 = (1 + al * ma) * (1 + al * mc) - ab2 * ab2;
        x = ((1 + al * mc) * n1 - ab2 * n2) / d2; y = ((1 + al * ma) * n2 - ab2 * n1) / d2;
        ll = x * x + y * y;
        if (!(ll > limit * limit)) break;
      }
      const double scale = limit / dsqrt(ll);
      n1 = x * scale; n2 = y * scale;
      sliding = 1;
    }
    t1_out = n1; t2_out = n2;
    const double d1 = n1 - o1, d2 = n2 - o2;
    const double px = t1x * d1 + t2x * d2, py = t1y * d1 + t2y * d2, pz = t1z * d1 + t2z * d2;
    s.vax -= ima * px; s.vay -= ima * py; s.vaz -= ima * pz;
    s.vbx += imb * px; s.vby += imb * py; s.vbz += imb * pz;
    s.wax -= f[6] * d1 + g[6] * d2; s.way -= f[7] * d1 + g[7] * d2; s.waz -= f[8] * d1 + g[8] * d2;
    s.wbx += f[9] * d1 + g[9] * d2; s.wby += f[10] * d1 + g[10] * d2; s.wbz += f[11] * d1 + g[11] * d2;
  }
  if (N <= 2) {
    const double most = patch * pressed;
    const double ux = s.wbx - s.wax, uy = s.wby - s.way, uz = s.wbz - s.waz;
    const double r1 = r1_in, r2 = r2_in;
    double n1 = r1 - rm1 * (ux * t1x + uy * t1y + uz * t1z);
    double n2 = r2 - rm2 * (ux * t2x + uy * t2y + uz * t2z);
    const double l2 = n1 * n1 + n2 * n2;
    if (l2 > most * most) {
      const double scale = most / dsqrt(l2);
      n1 *= scale; n2 *= scale;
    }
    r1_out = n1; r2_out = n2;
    const double d1 = n1 - r1, d2 = n2 - r2;
    const double lx = t1x * d1 + t2x * d2, ly = t1y * d1 + t2y * d2, lz = t1z * d1 + t2z * d2;
    s.wax -= Ia[0] * lx + Ia[1] * ly + Ia[2] * lz;
    s.way -= Ia[1] * lx + Ia[3] * ly + Ia[4] * lz;
    s.waz -= Ia[2] * lx + Ia[4] * ly + Ia[5] * lz;
    s.wbx += Ib[0] * lx + Ib[1] * ly + Ib[2] * lz;
    s.wby += Ib[1] * lx + Ib[3] * ly + Ib[4] * lz;
    s.wbz += Ib[2] * lx + Ib[4] * ly + Ib[5] * lz;
  }
  {
    const double limit = N == 1 ? SPIN_DISC * mu * patch * pressed : mu * twisted;
    const double s0 = s0_in;
    double s1 = s0 - spin_mass * ((s.wbx - s.wax) * nx + (s.wby - s.way) * ny + (s.wbz - s.waz) * nz);
    if (s1 > limit) s1 = limit; else if (s1 < -limit) s1 = -limit;
    s_out = s1;
    const double ds = s1 - s0, lx = nx * ds, ly = ny * ds, lz = nz * ds;
    s.wax -= Ia[0] * lx + Ia[1] * ly + Ia[2] * lz;
    s.way -= Ia[1] * lx + Ia[3] * ly + Ia[4] * lz;
    s.waz -= Ia[2] * lx + Ia[4] * ly + Ia[5] * lz;
    s.wbx += Ib[0] * lx + Ib[1] * ly + Ib[2] * lz;
    s.wby += Ib[1] * lx + Ib[3] * ly + Ib[4] * lz;
    s.wbz += Ib[2] * lx + Ib[4] * ly + Ib[5] * lz;
  }
  scatter(vw, sa, sb, ima, imb, s);
  // The row's results, written last.
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int i = m * P + k;
    if (!known && !bias) d.now[i] = now[k];
    d.normal_impulse[i] = impulse[k];
    if ((first >> k) & 1) d.relative_velocity[i] = closing_now[k];
    if ((larger >> k) & 1) d.max_normal_impulse[i] = pushed[k];
  }
  d.gap[m] = least;
  d.sliding[m] = sliding;
  d.tangent_impulse1[m] = t1_out;
  d.tangent_impulse2[m] = t2_out;
  if (N <= 2) {
    d.roll_impulse1[m] = r1_out;
    d.roll_impulse2[m] = r2_out;
  }
  d.spin_impulse[m] = s_out;
}

// src/world.js:1495-1530 (_accelerate) for one awake body.
__device__ __forceinline__ void accelerate_body(const SoftDev& d, const VW& vw, int g, int slot, double h, int substeps,
                                const double* grav) {
  if (d.type[g] != DYNAMIC || d.frozen[g] != 0) return;
  const double gx = grav[0] * h, gy = grav[1] * h, gz = grav[2] * h;
  const double most = SPIN_TURN / (h * substeps);
  double *v = vw.v(slot), *w = vw.w(slot);
  const double sc = d.gravity_scale[g];
  v[0] += gx * sc; v[1] += gy * sc; v[2] += gz * sc;
  if (d.damping[g] != 0) {
    const double k = 1 / (1 + h * d.damping[g]);
    v[0] *= k; v[1] *= k; v[2] *= k;
  }
  if (d.angular_damping[g] != 0) {
    const double k = 1 / (1 + h * d.angular_damping[g]);
    w[0] *= k; w[1] *= k; w[2] *= k;
  }
  if (d.shape[g] != SPHERE_SHAPE || d.framed[g] == 1) {
    const double w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    if (w2 > most * most) {
      const double k = most / dsqrt(w2);
      w[0] *= k; w[1] *= k; w[2] *= k;
    }
  }
}

// src/world.js:2820-2828 (advance): normalize(q + h (w, 0) q), returned.
struct Q4 {
  double x, y, z, w;
};
__device__ __forceinline__ Q4 advance(const Q4& q, double wx, double wy, double wz, double h) {
  const double nx = q.x + h * (wx * q.w + wy * q.z - wz * q.y);
  const double ny = q.y + h * (wy * q.w + wz * q.x - wx * q.z);
  const double nz = q.z + h * (wz * q.w + wx * q.y - wy * q.x);
  const double nw = q.w - h * (wx * q.x + wy * q.y + wz * q.z);
  const double l = 1 / dsqrt(nx * nx + ny * ny + nz * nz + nw * nw);
  return Q4{nx * l, ny * l, nz * l, nw * l};
}

// src/world.js:2831-2834 (angular): R I^-1 R^T L.
__device__ __forceinline__ V3 angular_of(const Q4& q, const double* inv, double lx, double ly, double lz) {
  const V3 b = rotate(q.x, q.y, q.z, q.w, lx, ly, lz, true);
  return rotate(q.x, q.y, q.z, q.w, b.x * inv[0], b.y * inv[1], b.z * inv[2], false);
}

// src/world.js:2777-2805 (untwist)
__device__ __forceinline__ Q4 untwist(const Q4& q, const double* inv, double lx, double ly, double lz, double most) {
  const V3 B = rotate(q.x, q.y, q.z, q.w, lx, ly, lz, true);
  const double bx = B.x, by = B.y, bz = B.z, ll = bx * bx + by * by + bz * bz;
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
    const double back = dsqrt(ll / (x * x + y * y + z * z));
    x *= back; y *= back; z *= back;
  }
  const double cx = y * bz - z * by, cy = z * bx - x * bz, cz = x * by - y * bx;
  const double cw = ll + x * bx + y * by + z * bz;
  const double l = dsqrt(cx * cx + cy * cy + cz * cz + cw * cw);
  if (!(l > 0)) return q;
  const double tx = cx / l, ty = cy / l, tz = cz / l, tw = cw / l;
  // q (x) T, src/core/math.js:231-238
  const double ox = q.w * tx + q.x * tw + q.y * tz - q.z * ty;
  const double oy = q.w * ty + q.y * tw + q.z * tx - q.x * tz;
  const double oz = q.w * tz + q.z * tw + q.x * ty - q.y * tx;
  const double ow = q.w * tw - q.x * tx - q.y * ty - q.z * tz;
  const double n = 1 / dsqrt(ox * ox + oy * oy + oz * oz + ow * ow);
  return Q4{ox * n, oy * n, oz * n, ow * n};
}

// src/world.js:2739-2763 (tumble): q and w of body g on by dt, carrying its angular momentum.
__device__ __forceinline__ void tumble(const SoftDev& d, double* w, int g, double dt) {
  double* qp = d.rotation + 4 * g;
  const double* I = d.inertia + 3 * g;
  const double* inv = d.inv_inertia + 3 * g;
  double* left = d.tumbled + 3 * g;
  Q4 q{qp[0], qp[1], qp[2], qp[3]};
  const double w0 = w[0], w1 = w[1], w2 = w[2];
  const V3 b = rotate(q.x, q.y, q.z, q.w, w0, w1, w2, true);
  const V3 L = rotate(q.x, q.y, q.z, q.w, b.x * I[0], b.y * I[1], b.z * I[2], false);
  const double lx = L.x, ly = L.y, lz = L.z;
  if (w0 != left[0] || w1 != left[1] || w2 != left[2]) d.ceiling[g] = w0 * lx + w1 * ly + w2 * lz;
  const double most = d.ceiling[g];
  const Q4 q4 = advance(q, w0, w1, w2, dt * 0.25);
  V3 v = angular_of(q4, inv, lx, ly, lz);
  q = advance(q, v.x, v.y, v.z, dt * 0.5);
  v = angular_of(q, inv, lx, ly, lz);
  if (v.x * lx + v.y * ly + v.z * lz > most * (1 + ROUNDED)) {
    q = untwist(q, inv, lx, ly, lz, most);
    v = angular_of(q, inv, lx, ly, lz);
  }
  qp[0] = q.x; qp[1] = q.y; qp[2] = q.z; qp[3] = q.w;
  w[0] = v.x; w[1] = v.y; w[2] = v.z;
  left[0] = v.x; left[1] = v.y; left[2] = v.z;
}

// src/world.js:1533-1563 (_move) for one awake body.
__device__ __forceinline__ void move_body(const SoftDev& d, const VW& vw, int g, int slot, double h) {
  if (d.frozen[g] != 0) return;
  double *p = d.position + 3 * g, *v = vw.v(slot), *w = vw.w(slot);
  p[0] += v[0] * h; p[1] += v[1] * h; p[2] += v[2] * h;
  const double wx = w[0], wy = w[1], wz = w[2];
  const bool turning = wx != 0 || wy != 0 || wz != 0;
  const double* I = d.inertia + 3 * g;
  if (turning && d.type[g] == DYNAMIC && d.lock[g] == 0 && d.joint_head[g] == 0 && (I[0] != I[1] || I[1] != I[2])) {
    tumble(d, w, g, h);
  } else if (turning) {
    double* qp = d.rotation + 4 * g;
    const Q4 q = advance(Q4{qp[0], qp[1], qp[2], qp[3]}, wx, wy, wz, h * 0.5);
    qp[0] = q.x; qp[1] = q.y; qp[2] = q.z; qp[3] = q.w;
  }
}

// src/contact.js:2229-2251 (restitution) for one world, in the order narrow() found
// the manifolds. Bounces on a shared body follow each other, so one lane applies them
// in that order; but which rows bounce at all (restitution, a point that pushed and
// met faster than the cut-off) doesn't change while they are applied, so the whole
// workgroup finds them first, a chunk of rows at a time, and compacts them in order
// into LDS for that lane. At rest almost no row bounces.
constexpr int BOUNCE_CHUNK = 512;

__device__ __forceinline__ bool bounces(const SoftDev& d, int m, int lm, int solved, double gravity) {
  const double e = d.restitution[m];
  if (e == 0 || lm >= solved) return false;
  const double threshold = dsqrt(2 * gravity * d.tolerance[m]);
  for (int k = 0; k < d.points[m]; k++) {
    const int i = m * P + k;
    if (!(d.relative_velocity[i] > -threshold || d.max_normal_impulse[i] == 0)) return true;
  }
  return false;
}

__device__ __forceinline__ void bounce_row(const SoftDev& d, const VW& vw, int m, int boff, double gravity) {
  const double e = d.restitution[m];
  const double threshold = dsqrt(2 * gravity * d.tolerance[m]);
  const int sa = d.a[m], sb = d.b[m], ga = boff + sa, gb = boff + sb;
  const int m3 = 3 * m;
  const double nx = d.normal[m3], ny = d.normal[m3 + 1], nz = d.normal[m3 + 2];
  const double ima = d.inv_mass[ga], imb = d.inv_mass[gb];
  for (int k = 0; k < d.points[m]; k++) {
    const int i = m * P + k;
    if (d.relative_velocity[i] > -threshold || d.max_normal_impulse[i] == 0) continue;
    const double* jac = d.jac + J * i;
    double *va = vw.v(sa), *wa = vw.w(sa), *vb = vw.v(sb), *wb = vw.w(sb);
    const double vn = (vb[0] - va[0]) * nx + (vb[1] - va[1]) * ny + (vb[2] - va[2]) * nz
        + wb[0] * jac[3] + wb[1] * jac[4] + wb[2] * jac[5] - wa[0] * jac[0] - wa[1] * jac[1] - wa[2] * jac[2];
    const double old = d.normal_impulse[i];
    double next = old - d.normal_mass[i] * (vn + e * d.relative_velocity[i]);
    if (next < 0) next = 0;
    const double l = next - old;
    d.normal_impulse[i] = next;
    // push(), src/contact.js:1743-1749: every component written, as the reference does.
    const double px = nx * l, py = ny * l, pz = nz * l;
    va[0] -= ima * px; va[1] -= ima * py; va[2] -= ima * pz;
    vb[0] += imb * px; vb[1] += imb * py; vb[2] += imb * pz;
    wa[0] -= jac[6] * l; wa[1] -= jac[7] * l; wa[2] -= jac[8] * l;
    wb[0] += jac[9] * l; wb[1] += jac[10] * l; wb[2] += jac[11] * l;
  }
}

__device__ __forceinline__ void restitution_world(const SoftDev& d, const VW& vw, int moff, int boff, int count,
                                                  int solved, double gravity, int* list, int* waves) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
  for (int base = 0; base < count; base += BOUNCE_CHUNK) {
    const int end = base + BOUNCE_CHUNK < count ? base + BOUNCE_CHUNK : count;
    // Each thread looks at consecutive rows f of the chunk; a wave's hits are counted by ballot.
    int total = 0;
    for (int f0 = base; f0 < end; f0 += blockDim.x) {
      const int f = f0 + (int)threadIdx.x;
      int m = -1;
      bool hit = false;
      if (f < end) {
        const int lm = d.row[moff + f];
        m = moff + lm;
        hit = bounces(d, m, lm, solved, gravity);
      }
      const uint64_t mask = __ballot(hit);
      if (lane == 0) waves[wave] = __popcll(mask);
      __syncthreads();
      int before = total;
      for (int k = 0; k < wave; k++) before += waves[k];
      if (hit) list[before + __popcll(mask & ((1ull << lane) - 1))] = m;
      for (int k = 0; k < nwaves; k++) total += waves[k];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      for (int k = 0; k < total; k++) bounce_row(d, vw, list[k], boff, gravity);
    }
    __syncthreads();
  }
}

// src/contact.js:1168-1182 (turned): body g's move over the step and its turn vector.
__device__ __forceinline__ void turned(const SoftDev& d, int g, double* T) {
  const double *p = d.position + 3 * g, *st = d.start + 3 * g, *q = d.rotation + 4 * g, *q0 = d.start_q + 4 * g;
  const double ax = q[0], ay = q[1], az = q[2], aw = q[3];
  const double bx = -q0[0], by = -q0[1], bz = -q0[2], bw = q0[3];
  double ux = aw * bx + ax * bw + ay * bz - az * by;
  double uy = aw * by + ay * bw + az * bx - ax * bz;
  double uz = aw * bz + az * bw + ax * by - ay * bx;
  double uw = aw * bw - ax * bx - ay * by - az * bz;
  if (uw < 0) { ux = -ux; uy = -uy; uz = -uz; uw = -uw; }
  const double sine = dsqrt(ux * ux + uy * uy + uz * uz);
  const double k = sine == 0 ? 0 : 2 * js_atan2(sine, uw) / sine;
  T[0] = p[0] - st[0]; T[1] = p[1] - st[1]; T[2] = p[2] - st[2];
  T[3] = ux * k; T[4] = uy * k; T[5] = uz * k;
}

// src/contact.js:1122-1153 (slipped) then 1091-1115 (keep) for row m, into pair gp.
__device__ __forceinline__ void store_row(const SoftDev& d, int m, int ga, int gb) {
  const int m3 = 3 * m, count = d.points[m];
  double pressed = 0;
  bool impact = false;
  for (int k = 0; k < count; k++) {
    const int i = m * P + k, i3 = 3 * i;
    pressed += d.normal_impulse[i];
    if (d.shift[i3] != 0 || d.shift[i3 + 1] != 0 || d.shift[i3 + 2] != 0) impact = true;
  }
  const double t1x = d.t1[m3], t1y = d.t1[m3 + 1], t1z = d.t1[m3 + 2];
  const double t2x = d.t2[m3], t2y = d.t2[m3 + 1], t2z = d.t2[m3 + 2];
  if (d.sliding[m] != 0 || pressed == 0 || impact) {
    d.drift1[m] = 0; d.drift2[m] = 0;
  } else {
    double T[6];
    double dx = 0, dy = 0, dz = 0;
    if (d.awake[gb] != 0) {
      turned(d, gb, T);
      const double rx = d.anchor_b[m3], ry = d.anchor_b[m3 + 1], rz = d.anchor_b[m3 + 2];
      dx = T[0] + T[4] * rz - T[5] * ry;
      dy = T[1] + T[5] * rx - T[3] * rz;
      dz = T[2] + T[3] * ry - T[4] * rx;
    }
    if (d.awake[ga] != 0) {
      turned(d, ga, T);
      const double rx = d.anchor_a[m3], ry = d.anchor_a[m3 + 1], rz = d.anchor_a[m3 + 2];
      dx -= T[0] + T[4] * rz - T[5] * ry;
      dy -= T[1] + T[5] * rx - T[3] * rz;
      dz -= T[2] + T[3] * ry - T[4] * rx;
    }
    d.drift1[m] += dx * t1x + dy * t1y + dz * t1z;
    d.drift2[m] += dx * t2x + dy * t2y + dz * t2z;
  }
  // A level pair's manifolds go into its Memory, which the host keeps. What the row stores
  // for its pair is kept by row here; the download writes it into the pair's columns.
  if (d.sub[m] >= 0) return;
  const int k3 = 3 * m;
  d.pair_points[m] = (uint8_t)count;
  const double r1 = d.roll_impulse1[m], r2 = d.roll_impulse2[m], sp = d.spin_impulse[m];
  d.pair_twist[k3] = t1x * r1 + t2x * r2 + d.normal[m3] * sp;
  d.pair_twist[k3 + 1] = t1y * r1 + t2y * r2 + d.normal[m3 + 1] * sp;
  d.pair_twist[k3 + 2] = t1z * r1 + t2z * r2 + d.normal[m3 + 2] * sp;
  const double l1 = d.tangent_impulse1[m], l2 = d.tangent_impulse2[m];
  d.pair_friction[k3] = t1x * l1 + t2x * l2;
  d.pair_friction[k3 + 1] = t1y * l1 + t2y * l2;
  d.pair_friction[k3 + 2] = t1z * l1 + t2z * l2;
  const double d1 = d.drift1[m], d2 = d.drift2[m];
  d.pair_slip[k3] = t1x * d1 + t2x * d2;
  d.pair_slip[k3 + 1] = t1y * d1 + t2y * d2;
  d.pair_slip[k3 + 2] = t1z * d1 + t2z * d2;
  for (int k = 0; k < count; k++) {
    d.pair_pid[P * m + k] = d.pid[m * P + k];
    d.pair_normal_impulse[P * m + k] = d.normal_impulse[m * P + k];
  }
}

// One row of a sweep: warm start (KIND 0), solve (1) or relax (2), by its point count.
template <int KIND>
__device__ __forceinline__ void row_of(const SoftDev& d, const VW& vw, const Soft& soft, double h, bool known, int m,
                                       int boff) {
  const int sa = d.a[m], sb = d.b[m], ga = boff + sa, gb = boff + sb;
  switch (d.points[m]) {
    case 1:
      if (KIND == 0) warm_row<1>(d, vw, m, ga, gb, sa, sb); else solve_row<1>(d, vw, soft, h, KIND == 1, known, m, ga, gb, sa, sb);
      break;
    case 2:
      if (KIND == 0) warm_row<2>(d, vw, m, ga, gb, sa, sb); else solve_row<2>(d, vw, soft, h, KIND == 1, known, m, ga, gb, sa, sb);
      break;
    case 3:
      if (KIND == 0) warm_row<3>(d, vw, m, ga, gb, sa, sb); else solve_row<3>(d, vw, soft, h, KIND == 1, known, m, ga, gb, sa, sb);
      break;
    default:
      if (KIND == 0) warm_row<4>(d, vw, m, ga, gb, sa, sb); else solve_row<4>(d, vw, soft, h, KIND == 1, known, m, ga, gb, sa, sb);
      break;
  }
}

// One sweep over a world's colours, a barrier after each, then its overflow in order.
template <int KIND>
__device__ __forceinline__ void sweep(const SoftDev& d, const VW& vw, const Soft& soft, double h, bool known,
                                      int moff, int boff, const int* starts) {
  for (int c = 0; c < COLOURS; c++) {
    const int from = starts[c], to = starts[c + 1];
    if (from == to) continue;
    for (int lm = from + (int)threadIdx.x; lm < to; lm += blockDim.x) row_of<KIND>(d, vw, soft, h, known, moff + lm, boff);
    __syncthreads();
  }
  const int from = starts[OVERFLOW], to = starts[OVERFLOW + 1];
  if (from < to) {
    if (threadIdx.x == 0) {
      for (int lm = from; lm < to; lm++) row_of<KIND>(d, vw, soft, h, known, moff + lm, boff);
    }
    __syncthreads();
  }
}

// src/world.js:961-1031 for the worlds blockIdx.x, blockIdx.x + gridDim.x, ...
// RESIDENT: the world's body velocities are copied into LDS for the phase and back
// after it (the host chooses it when every world's bodies fit the LDS it gives).
template <bool RESIDENT, int THREADS, int WAVES>
__global__ void __launch_bounds__(THREADS, WAVES) soft_step_kernel(SoftDev d) {
  __shared__ int starts[BM_STARTS_DEV];
  __shared__ int bounce_list[BOUNCE_CHUNK];
  __shared__ int bounce_waves[THREADS / 64];
  extern __shared__ double lds[];
  for (int at = blockIdx.x; at < d.n_worlds; at += gridDim.x) {
    const int wld = d.world_order ? d.world_order[at] : at;
    const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
    const int nbodies = d.body_offset[wld + 1] - boff;
    const int aoff = d.awake_offset[wld], awake = d.awake_offset[wld + 1] - aoff;
    const int count = d.manifold_offset[wld + 1] - moff, solved = d.solved[wld];
    const int substeps = d.substeps[wld];
    const double h = d.dt[wld] / substeps;
    const double* grav = d.gravity + 3 * wld;
    __syncthreads();
    for (int k = threadIdx.x; k < BM_STARTS_DEV; k += blockDim.x) starts[k] = d.starts[BM_STARTS_DEV * wld + k];
    const bool touching = count > 0;
    for (int i = threadIdx.x; i < awake; i += blockDim.x) {
      const int g = boff + d.awake_list[aoff + i];
      if (touching) inertia_of(d, g);
      for (int k = 0; k < 3; k++) d.start[3 * g + k] = d.position[3 * g + k];
      for (int k = 0; k < 4; k++) d.start_q[4 * g + k] = d.rotation[4 * g + k];
    }
    VW vw;
    if (RESIDENT) {
      for (int i = threadIdx.x; i < nbodies; i += blockDim.x) {
        const double *v = d.velocity + 3 * (boff + i), *w = d.angular + 3 * (boff + i);
        double* o = lds + 6 * i;
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = w[0]; o[4] = w[1]; o[5] = w[2];
      }
      vw = VW{lds, lds + 3, 6};
    } else {
      vw = VW{d.velocity + 3 * boff, d.angular + 3 * boff, 3};
    }
    __syncthreads();
    Soft soft{0, 1, 0};
    if (touching) {
      for (int lm = threadIdx.x; lm < solved; lm += blockDim.x) {
        const int m = moff + lm, sa = d.a[m], sb = d.b[m];
        prepare_row(d, vw, m, boff + sa, boff + sb, sa, sb);
      }
      soft = make_soft(STIFFNESS / h, DAMPING_RATIO, h);
      __syncthreads();
    }
    for (int sub = 0; sub < substeps; sub++) {
      for (int i = threadIdx.x; i < awake; i += blockDim.x) {
        const int slot = d.awake_list[aoff + i];
        accelerate_body(d, vw, boff + slot, slot, h, substeps, grav);
      }
      __syncthreads();
      if (touching) {
        sweep<0>(d, vw, soft, h, false, moff, boff, starts);
        sweep<1>(d, vw, soft, h, sub > 0, moff, boff, starts);
      }
      for (int i = threadIdx.x; i < awake; i += blockDim.x) {
        const int slot = d.awake_list[aoff + i];
        move_body(d, vw, boff + slot, slot, h);
      }
      __syncthreads();
      if (touching) sweep<2>(d, vw, soft, h, false, moff, boff, starts);
    }
    if (touching) {
      restitution_world(d, vw, moff, boff, count, solved, dsqrt(grav[0] * grav[0] + grav[1] * grav[1] + grav[2] * grav[2]),
                        bounce_list, bounce_waves);
      for (int f = threadIdx.x; f < count; f += blockDim.x) {
        const int lm = d.row[moff + f];
        if (lm >= solved) continue;
        const int m = moff + lm;
        store_row(d, m, boff + d.a[m], boff + d.b[m]);
      }
    }
    if (RESIDENT) {
      for (int i = threadIdx.x; i < nbodies; i += blockDim.x) {
        double *v = d.velocity + 3 * (boff + i), *w = d.angular + 3 * (boff + i);
        const double* o = lds + 6 * i;
        v[0] = o[0]; v[1] = o[1]; v[2] = o[2]; w[0] = o[3]; w[1] = o[4]; w[2] = o[5];
      }
    }
  }
}

// The most dynamic LDS a workgroup of the resident kernel may take on `device`:
// the device's LDS a workgroup less the kernel's own static LDS.
size_t soft_step_lds_limit(int device) {
  int most = 0;
  if (hipDeviceGetAttribute(&most, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return 0;
  const size_t own = (BM_STARTS_DEV + BOUNCE_CHUNK + 8) * sizeof(int) + 64;
  return (size_t)most > own ? (size_t)most - own : 0;
}

// A barrier over every workgroup of a cooperative launch, in two levels: a workgroup's first
// thread arrives at its group's counter (workgroup i in group i % 8: workgroups are handed to
// the eight XCDs in turn, so a group's arrivals come from one XCD), the last of a group at
// the top counter, and the last there opens the next generation, which every other first
// thread waits for. Each counter is reset by its last arrival before it goes on, so all are
// back at 0 when the next generation opens. Every arrival releases at agent scope, a group's
// last arrival acquires before it arrives at the top, and a waiter polls without ordering
// (an acquire invalidates the XCD's L2, which the other workgroups' rows are reading) and
// acquires once when the generation changes: every write before the barrier is seen by every
// read after it, as with cooperative_groups' grid.sync() (which took 32-40 us a barrier over
// 256 workgroups on the MI355X).
// words: the generation at [0], group g's counter at [WIDE_STRIDE * (g + 1)], the top's at
// [WIDE_STRIDE * (WIDE_GROUPS + 1)]; every counter 0 before the first launch.
struct WideBar {
  unsigned* words;
  __device__ void sync() const {
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned gen = __hip_atomic_load(&words[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned g = blockIdx.x % WIDE_GROUPS;
      const unsigned members = (gridDim.x - g + WIDE_GROUPS - 1) / WIDE_GROUPS;
      const unsigned groups = gridDim.x < WIDE_GROUPS ? gridDim.x : WIDE_GROUPS;
      unsigned* group = &words[WIDE_STRIDE * (g + 1)];
      unsigned* top = &words[WIDE_STRIDE * (WIDE_GROUPS + 1)];
      bool last = false;
      if (__hip_atomic_fetch_add(group, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(group, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == groups - 1) {
          __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_add(&words[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          last = true;
        }
      }
      if (!last) {
        while (__hip_atomic_load(&words[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
          __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
    }
    __syncthreads();
  }
};

// One sweep over the world's colours across the whole grid: each colour's manifolds spread
// over every thread of every workgroup, a grid barrier after each; the overflow in order by
// one thread.
// Which path through row_of a solved row takes: the instance for its point count.
__device__ __forceinline__ int path_of(const SoftDev& d, int m) {
  const int n = d.points[m];
  return n < 1 ? 0 : n > 4 ? 3 : n - 1;
}
constexpr int PATHS = 4;

// Colour c's rows grouped by path into pick[from .. to), each group in row order: a wave then
// takes one path through row_of instead of each of its rows' paths one after another. The rows
// of a colour share no moving body, so the order threads take them in changes no result.
// Workgroup-wide, stable: a chunk of THREADS rows at a time, each path's rows placed by a
// ballot within the wave and a prefix over the waves.
template <int THREADS>
__device__ __forceinline__ void group_colour(const SoftDev& d, int moff, int from, int to, int* pick) {
  __shared__ int counts[THREADS / 64], cursor[PATHS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x < PATHS) cursor[threadIdx.x] = 0;
  __syncthreads();
  // First the size of each path, so each group starts where the ones before it end.
  for (int s = from; s < to; s += THREADS) {
    const int i = s + (int)threadIdx.x;
    const int path = i < to ? path_of(d, moff + i) : -1;
    for (int k = 0; k < PATHS; k++) {
      const unsigned long long mask = __ballot(path == k);
      if (lane == 0 && mask) atomicAdd(&cursor[k], __popcll(mask));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int at = from;
    for (int k = 0; k < PATHS; k++) {
      const int n = cursor[k];
      cursor[k] = at;
      at += n;
    }
  }
  __syncthreads();
  for (int s = from; s < to; s += THREADS) {
    const int i = s + (int)threadIdx.x;
    const int path = i < to ? path_of(d, moff + i) : -1;
    for (int k = 0; k < PATHS; k++) {
      const unsigned long long mask = __ballot(path == k);
      if (lane == 0) counts[wave] = __popcll(mask);
      __syncthreads();
      int before = cursor[k];
      for (int w = 0; w < wave; w++) before += counts[w];
      if (path == k) pick[moff + before + __popcll(mask & ((1ull << lane) - 1))] = i;
      __syncthreads();
      if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < THREADS / 64; w++) total += counts[w];
        cursor[k] += total;
      }
      __syncthreads();
    }
  }
}

// World wld's colours grouped for the grid-wide soft step, workgroup c taking colour c.
template <int THREADS>
__global__ void __launch_bounds__(THREADS) group_kernel(SoftDev d, int wld, int* pick) {
  const int moff = d.manifold_offset[wld];
  const int c = blockIdx.x;
  const int from = d.starts[BM_STARTS_DEV * wld + c], to = d.starts[BM_STARTS_DEV * wld + c + 1];
  if (from < to) group_colour<THREADS>(d, moff, from, to, pick);
}

template <int KIND>
__device__ __forceinline__ void sweep_wide(const WideBar& grid, const SoftDev& d, const VW& vw,
                                           const Soft& soft, double h, bool known, int moff, int boff,
                                           const int* starts, const int* pick, int tid, int nth) {
  // A colour's rows shared out in equal runs, one run a workgroup, not in blocks of one
  // workgroup's threads: a colour is a few thousand rows, and spread over every CU their loads'
  // addresses go through every CU's memory pipeline rather than queueing in a dozen; a run's
  // rows are neighbours in the grouped order, so a field of a few bytes a row is read a cache
  // line for several lanes at once.
  for (int c = 0; c < COLOURS; c++) {
    const int from = starts[c], to = starts[c + 1];
    if (from == to) continue;
    const int run = (to - from + (int)gridDim.x - 1) / (int)gridDim.x;
    const int first = from + (int)blockIdx.x * run, last = first + run < to ? first + run : to;
    for (int j = first + (int)threadIdx.x; j < last; j += (int)blockDim.x)
      row_of<KIND>(d, vw, soft, h, known, moff + pick[moff + j], boff);
    grid.sync();
  }
  const int from = starts[OVERFLOW], to = starts[OVERFLOW + 1];
  if (from < to) {
    if (tid == 0) {
      for (int lm = from; lm < to; lm++) row_of<KIND>(d, vw, soft, h, known, moff + lm, boff);
    }
    grid.sync();
  }
}

// World `wld`'s soft step across the whole grid, for a world too large for one workgroup to
// sweep quickly (tens of thousands of manifolds): the phases of soft_step_kernel, each one's
// bodies or manifolds spread over every thread, a grid barrier where that kernel has a
// workgroup barrier (a cooperative launch: every workgroup is resident, so the barrier is
// reached by all). Manifolds of one colour share no moving body, so the order threads take
// them in changes no result; the overflow and the bounces stay serial in the reference's
// order. Velocities stay in HBM.
template <int THREADS>
__global__ void __launch_bounds__(THREADS, 1) soft_wide_kernel(SoftDev d, int wld, unsigned* words, int* pick) {
  const WideBar grid{words};
  __shared__ int starts[BM_STARTS_DEV];
  __shared__ int bounce_list[BOUNCE_CHUNK];
  __shared__ int bounce_waves[THREADS / 64];
  const int tid = blockIdx.x * THREADS + threadIdx.x, nth = gridDim.x * THREADS;
  const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
  const int aoff = d.awake_offset[wld], awake = d.awake_offset[wld + 1] - aoff;
  const int count = d.manifold_offset[wld + 1] - moff, solved = d.solved[wld];
  const int substeps = d.substeps[wld];
  const double h = d.dt[wld] / substeps;
  const double* grav = d.gravity + 3 * wld;
  for (int k = threadIdx.x; k < BM_STARTS_DEV; k += THREADS) starts[k] = d.starts[BM_STARTS_DEV * wld + k];
  const bool touching = count > 0;
  for (int i = tid; i < awake; i += nth) {
    const int g = boff + d.awake_list[aoff + i];
    if (touching) inertia_of(d, g);
    for (int k = 0; k < 3; k++) d.start[3 * g + k] = d.position[3 * g + k];
    for (int k = 0; k < 4; k++) d.start_q[4 * g + k] = d.rotation[4 * g + k];
  }
  const VW vw{d.velocity + 3 * boff, d.angular + 3 * boff, 3};
  grid.sync();
  Soft soft{0, 1, 0};
  if (touching) {
    for (int lm = tid; lm < solved; lm += nth) {
      const int m = moff + lm, sa = d.a[m], sb = d.b[m];
      prepare_row(d, vw, m, boff + sa, boff + sb, sa, sb);
    }
    soft = make_soft(STIFFNESS / h, DAMPING_RATIO, h);
    grid.sync();
  }
  for (int sub = 0; sub < substeps; sub++) {
    for (int i = tid; i < awake; i += nth) {
      const int slot = d.awake_list[aoff + i];
      accelerate_body(d, vw, boff + slot, slot, h, substeps, grav);
    }
    grid.sync();
    if (touching) {
      sweep_wide<0>(grid, d, vw, soft, h, false, moff, boff, starts, pick, tid, nth);
      sweep_wide<1>(grid, d, vw, soft, h, sub > 0, moff, boff, starts, pick, tid, nth);
    }
    for (int i = tid; i < awake; i += nth) {
      const int slot = d.awake_list[aoff + i];
      move_body(d, vw, boff + slot, slot, h);
    }
    grid.sync();
    if (touching) sweep_wide<2>(grid, d, vw, soft, h, false, moff, boff, starts, pick, tid, nth);
  }
  if (touching) {
    if (blockIdx.x == 0)
      restitution_world(d, vw, moff, boff, count, solved, dsqrt(grav[0] * grav[0] + grav[1] * grav[1] + grav[2] * grav[2]),
                        bounce_list, bounce_waves);
    grid.sync();
    for (int f = tid; f < count; f += nth) {
      const int lm = d.row[moff + f];
      if (lm >= solved) continue;
      const int m = moff + lm;
      store_row(d, m, boff + d.a[m], boff + d.b[m]);
    }
  }
}

// The grid-wide soft step of world `wld`: a thread for every second unit of its work (a
// colour holds at most a fifth to a third of a world's manifolds, so a colour's rows still
// take a thread each, and a barrier costs less the fewer workgroups arrive at it), as many
// workgroups as can be resident at most.
hipError_t launch_soft_wide(const SoftDev& d, int wld, int work, unsigned* words, int* pick, int device,
                            hipStream_t stream, int* blocks_out) {
  constexpr int T = 256;
  int per_cu = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, soft_wide_kernel<T>, T, 0);
  if (e != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) return e;
  int blocks = per_cu * cus;
  const int wanted = (work + 2 * T - 1) / (2 * T);
  if (wanted < blocks) blocks = wanted;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(group_kernel<T>, dim3(COLOURS), dim3(T), 0, stream, d, wld, pick);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  SoftDev arg = d;
  void* args[] = {&arg, &wld, &words, &pick};
  e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(soft_wide_kernel<T>), dim3(blocks), dim3(T), args, 0,
                                 stream);
  *blocks_out = e == hipSuccess ? blocks : 0;
  return e;
}

// Launch shapes: threads a workgroup, and registers sized for 1 or 2 waves a SIMD.
void launch_soft_step(const SoftDev& d, int grid, int shape, size_t lds_bytes, hipStream_t stream) {
  const dim3 g(grid);
  if (lds_bytes > 0) {
    if (shape == 1) hipLaunchKernelGGL((soft_step_kernel<true, 256, 2>), g, dim3(256), lds_bytes, stream, d);
    else if (shape == 2) hipLaunchKernelGGL((soft_step_kernel<true, 512, 2>), g, dim3(512), lds_bytes, stream, d);
    else hipLaunchKernelGGL((soft_step_kernel<true, 256, 1>), g, dim3(256), lds_bytes, stream, d);
  } else {
    if (shape == 1) hipLaunchKernelGGL((soft_step_kernel<false, 256, 2>), g, dim3(256), 0, stream, d);
    else if (shape == 2) hipLaunchKernelGGL((soft_step_kernel<false, 512, 2>), g, dim3(512), 0, stream, d);
    else hipLaunchKernelGGL((soft_step_kernel<false, 256, 1>), g, dim3(256), 0, stream, d);
  }
}

}  // namespace bm
