// The adjoint form's arithmetic: a step of a world with its contacts held on their
// bodies, taped, and the step's reverse, for the device (adjoint.hip) and, in the
// tests, the host.
//
// What is differentiated. A world's trajectory is first stepped as the reference steps
// it; each step's solved manifolds are kept as REUSE keeps one (src/contact.js:388-508):
// each point's end on a in a's frame and its end on b in b's frame, the normal in a's
// frame, the feature ids. A held step places every manifold again from where its bodies
// are (the REUSE rebuild), carries the impulses the pair kept (src/contact.js:999-1036),
// and runs the soft step (src/world.js:961-1031: inertias, prepare, 8 substeps of
// accelerate / warm start / solve / move / relax, restitution, store) with the
// reference's arithmetic. Its derivative with respect to the initial velocities and the
// bodies' mass, friction and restitution is taken on the forward pass's active set (the
// Murty sets, the friction cone, roll and spin clamps, bounces, the tumble's branches):
// which manifolds a step has is part of that set, where their points lie moves with the
// bodies.
//
// How. The forward pass keeps each body's state every K steps and every step's stored
// impulses (what the pairs carry to the next step). The reverse pass restores a
// segment's first state, steps the segment again writing a tape (each row's velocities
// and impulses as each solve and relax sweep found them, each body's state before each
// substep and move), then walks the steps, substeps, sweeps and colours backwards: a
// row's reverse recomputes the row from its tape and pushes the adjoints of its outputs
// back onto its bodies' velocities, its impulses and what prepare derived (the lever
// terms, the block, the tangent and spin masses), which prepare's reverse then turns into
// adjoints of the bodies' poses, inertias and masses. The small nonlinear maps (the
// integrator's turn, tumble and untwist, the rebuild, the slip's turn angle) are
// differentiated by forward-mode dual numbers over their few inputs.
//
// Every function here is __host__ __device__; a world's step runs as one workgroup
// (Ctx: its threads and barrier) or, on the host, as one thread.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jsmath.h"

namespace bm {
namespace adj {

#define HD __host__ __device__ __forceinline__

constexpr int P = 4, J = 12, FJ = 24;
constexpr int COLOURS = 32, OVERFLOW = COLOURS, STARTS = COLOURS + 3;
constexpr int DYNAMIC = 0;
constexpr int SPHERE_SHAPE = 1, CAPSULE_SHAPE = 2;
constexpr double STIFFNESS = 1.0 / 8;               // src/world.js:115
constexpr double DAMPING_RATIO = 10;                // src/world.js:118
constexpr double SPIN_TURN = 3.141592653589793;     // src/world.js:171
constexpr double SPIN_DISC = 2.0 / 3;               // src/contact.js:146
constexpr double COMPLIANCE = 1.0 / 1048576;        // src/forest.js:67
constexpr double ROUNDED = 1.0 / 17592186044416.0;  // src/world.js:2809
constexpr double MATH_PI = 3.141592653589793;
// A solve or relax row's tape, field by field (each field a column over the working rows,
// so a colour's rows read and write it coalesced): what the row read (its bodies'
// velocities, its impulses, the largest pushes so far) and what its reverse needs of what
// it worked out (targets, closing speeds, the pushes, the friction cone's and the clamps'
// outcomes and the velocities between its stages).
enum : int {
  T_S = 0, T_IMP = 12, T_O1 = 16, T_O2 = 17, T_R1 = 18, T_R2 = 19, T_S0 = 20, T_MOST = 21, T_TARGET = 25,
  T_CLOSING = 29, T_RAW = 33, T_SOLVED = 37, T_NEXT = 41, T_DD = 45, T_FLAGS = 49, T_PRESSED = 50, T_TWISTED = 51,
  T_MID = 52, T_V1 = 64, T_V2 = 65, T_B1 = 66, T_B2 = 67, T_BB = 68, T_LIMIT = 69, T_AL = 70, T_X = 71, T_Y = 72,
  T_LL = 73, T_T1 = 74, T_T2 = 75, T_U = 76, T_RN1 = 79, T_RN2 = 80, T_RMOST = 81, T_R1OUT = 82, T_R2OUT = 83,
  T_DW = 84, T_SOUT = 87, ROW_TAPE = 88
};
struct TapeRow {
  double* p;
  size_t stride;
  HD double& operator[](int f) const { return p[(size_t)f * stride]; }
};
// Body state: p 3, q 4, v 3, w 3, left 3, ceiling 1.
constexpr int STATE = 17;

// ---------------------------------------------------------------------------------
// Forward-mode dual numbers, for the small nonlinear maps.
template <int W>
struct Dual {
  double v;
  double d[W];
};
template <int W>
HD Dual<W> dual(double x) {
  Dual<W> r;
  r.v = x;
  for (int k = 0; k < W; k++) r.d[k] = 0;
  return r;
}
template <int W>
HD Dual<W> operator+(const Dual<W>& a, const Dual<W>& b) {
  Dual<W> r;
  r.v = a.v + b.v;
  for (int k = 0; k < W; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int W>
HD Dual<W> operator-(const Dual<W>& a, const Dual<W>& b) {
  Dual<W> r;
  r.v = a.v - b.v;
  for (int k = 0; k < W; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int W>
HD Dual<W> operator-(const Dual<W>& a) {
  Dual<W> r;
  r.v = -a.v;
  for (int k = 0; k < W; k++) r.d[k] = -a.d[k];
  return r;
}
template <int W>
HD Dual<W> operator*(const Dual<W>& a, const Dual<W>& b) {
  Dual<W> r;
  r.v = a.v * b.v;
  for (int k = 0; k < W; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int W>
HD Dual<W> operator/(const Dual<W>& a, const Dual<W>& b) {
  Dual<W> r;
  r.v = a.v / b.v;
  const double ib = 1 / b.v;
  for (int k = 0; k < W; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int W>
HD Dual<W> operator+(const Dual<W>& a, double b) { Dual<W> r = a; r.v += b; return r; }
template <int W>
HD Dual<W> operator+(double a, const Dual<W>& b) { Dual<W> r = b; r.v += a; return r; }
template <int W>
HD Dual<W> operator-(const Dual<W>& a, double b) { Dual<W> r = a; r.v -= b; return r; }
template <int W>
HD Dual<W> operator-(double a, const Dual<W>& b) { Dual<W> r = -b; r.v += a; return r; }
template <int W>
HD Dual<W> operator*(const Dual<W>& a, double b) {
  Dual<W> r;
  r.v = a.v * b;
  for (int k = 0; k < W; k++) r.d[k] = a.d[k] * b;
  return r;
}
template <int W>
HD Dual<W> operator*(double a, const Dual<W>& b) { return b * a; }
template <int W>
HD Dual<W> operator/(const Dual<W>& a, double b) {
  Dual<W> r;
  r.v = a.v / b;
  for (int k = 0; k < W; k++) r.d[k] = a.d[k] / b;
  return r;
}
template <int W>
HD Dual<W> operator/(double a, const Dual<W>& b) { return dual<W>(a) / b; }
template <int W>
HD Dual<W> sqrt_of(const Dual<W>& a) {
  Dual<W> r;
  r.v = dsqrt(a.v);
  const double k = r.v > 0 ? 0.5 / r.v : 0;
  for (int i = 0; i < W; i++) r.d[i] = a.d[i] * k;
  return r;
}
HD double sqrt_of(double a) { return dsqrt(a); }
// atan2 of the reference (fdlibm), and its derivative (x dy - y dx) / (x^2 + y^2).
template <int W>
HD Dual<W> atan2_of(const Dual<W>& y, const Dual<W>& x) {
  Dual<W> r;
  r.v = js_atan2(y.v, x.v);
  const double n = x.v * x.v + y.v * y.v;
  for (int i = 0; i < W; i++) r.d[i] = n > 0 ? (x.v * y.d[i] - y.v * x.d[i]) / n : 0;
  return r;
}
HD double atan2_of(double y, double x) { return js_atan2(y, x); }
HD double val(double x) { return x; }
template <int W>
HD double val(const Dual<W>& x) { return x.v; }
// A +0 of x's kind, with no derivative.
HD double zero_of(double) { return 0; }
template <int W>
HD Dual<W> zero_of(const Dual<W>&) { return dual<W>(0); }

// The vector-Jacobian product of f: NI inputs -> NO outputs, at `in`, by forward-mode
// duals W directions at a time: in_bar += J^T out_bar.
// Directions a pass of the small maps' reverses carries (placement, slip turn, tumble, world
// inverse inertia): wider passes evaluate each map fewer times, narrower ones hold fewer
// registers in the backward kernel.
constexpr int DW = 2;
template <int W, int NI, int NO, class F>
HD void vjp(const double (&in)[NI], const double (&out_bar)[NO], double (&in_bar)[NI], F f) {
  for (int base = 0; base < NI; base += W) {
    Dual<W> x[NI], y[NO];
    for (int i = 0; i < NI; i++) {
      x[i] = dual<W>(in[i]);
      if (i >= base && i < base + W) x[i].d[i - base] = 1;
    }
    f(x, y);
    for (int i = base; i < NI && i < base + W; i++) {
      double s = 0;
      for (int o = 0; o < NO; o++) s += out_bar[o] * y[o].d[i - base];
      in_bar[i] += s;
    }
  }
}

// ---------------------------------------------------------------------------------
// Rotations and the integrator's turn, over any scalar (double or dual).

// src/core/math.js:211-220 (rotate), `inverse` by the conjugate.
template <class T>
HD void rotate_t(const T* q, const T* v, bool inverse, T* out) {
  const double s = inverse ? -1 : 1;
  const T ux = q[0] * s, uy = q[1] * s, uz = q[2] * s, w = q[3];
  const T tx = 2 * (uy * v[2] - uz * v[1]), ty = 2 * (uz * v[0] - ux * v[2]), tz = 2 * (ux * v[1] - uy * v[0]);
  out[0] = v[0] + w * tx + (uy * tz - uz * ty);
  out[1] = v[1] + w * ty + (uz * tx - ux * tz);
  out[2] = v[2] + w * tz + (ux * ty - uy * tx);
}

// src/core/math.js:223-228: the rotation's columns.
template <class T>
HD void axes_t(const T* q, T* ax) {
  const T x = q[0], y = q[1], z = q[2], w = q[3];
  ax[0] = 1 - 2 * (y * y + z * z); ax[1] = 2 * (x * y + z * w); ax[2] = 2 * (x * z - y * w);
  ax[3] = 2 * (x * y - z * w); ax[4] = 1 - 2 * (x * x + z * z); ax[5] = 2 * (y * z + x * w);
  ax[6] = 2 * (x * z + y * w); ax[7] = 2 * (y * z - x * w); ax[8] = 1 - 2 * (x * x + y * y);
}

// src/world.js:2820-2828 (advance): normalize(q + h (w, 0) q).
template <class T>
HD void advance_t(const T* q, const T* w, double h, T* out) {
  const T nx = q[0] + h * (w[0] * q[3] + w[1] * q[2] - w[2] * q[1]);
  const T ny = q[1] + h * (w[1] * q[3] + w[2] * q[0] - w[0] * q[2]);
  const T nz = q[2] + h * (w[2] * q[3] + w[0] * q[1] - w[1] * q[0]);
  const T nw = q[3] - h * (w[0] * q[0] + w[1] * q[1] + w[2] * q[2]);
  const T l = 1 / sqrt_of(nx * nx + ny * ny + nz * nz + nw * nw);
  out[0] = nx * l; out[1] = ny * l; out[2] = nz * l; out[3] = nw * l;
}

// src/world.js:2831-2834 (angular): R I^-1 R^T L.
template <class T>
HD void angular_t(const T* q, const T* inv, const T* L, T* out) {
  T b[3], c[3];
  rotate_t(q, L, true, b);
  c[0] = b[0] * inv[0]; c[1] = b[1] * inv[1]; c[2] = b[2] * inv[2];
  rotate_t(q, c, false, out);
}

// src/world.js:2777-2805 (untwist)
template <class T>
HD void untwist_t(const T* q, const T* inv, const T* L, const T& most, T* out) {
  T B[3];
  rotate_t(q, L, true, B);
  const T bx = B[0], by = B[1], bz = B[2], ll = bx * bx + by * by + bz * bz;
  T x = bx, y = by, z = bz;
  for (int round = 0; round < 2; round++) {
    const T gx = x * inv[0], gy = y * inv[1], gz = z * inv[2];
    const T over = x * gx + y * gy + z * gz - most;
    if (!(val(over) > 0)) break;
    const T k = (gx * x + gy * y + gz * z) / ll, tx = gx - k * x, ty = gy - k * y, tz = gz - k * z;
    const T tt = tx * tx + ty * ty + tz * tz;
    if (!(val(tt) > 0)) break;
    const T a = 0.5 * over / tt;
    x = x - a * tx; y = y - a * ty; z = z - a * tz;
    const T back = sqrt_of(ll / (x * x + y * y + z * z));
    x = x * back; y = y * back; z = z * back;
  }
  const T cx = y * bz - z * by, cy = z * bx - x * bz, cz = x * by - y * bx;
  const T cw = ll + x * bx + y * by + z * bz;
  const T l = sqrt_of(cx * cx + cy * cy + cz * cz + cw * cw);
  if (!(val(l) > 0)) {
    for (int k = 0; k < 4; k++) out[k] = q[k];
    return;
  }
  const T tx = cx / l, ty = cy / l, tz = cz / l, tw = cw / l;
  const T ox = q[3] * tx + q[0] * tw + q[1] * tz - q[2] * ty;
  const T oy = q[3] * ty + q[1] * tw + q[2] * tx - q[0] * tz;
  const T oz = q[3] * tz + q[2] * tw + q[0] * ty - q[1] * tx;
  const T ow = q[3] * tw - q[0] * tx - q[1] * ty - q[2] * tz;
  const T n = 1 / sqrt_of(ox * ox + oy * oy + oz * oz + ow * ow);
  out[0] = ox * n; out[1] = oy * n; out[2] = oz * n; out[3] = ow * n;
}

// src/world.js:2739-2763 (tumble). In: q 4, w 3, left 3, ceiling 1, inertia 3, inverse 3.
// Out: q 4, w 3 (= left), ceiling 1.
template <class T>
HD void tumble_t(const T* q0, const T* w, const T* left, const T& ceiling_in, const T* I, const T* inv, double dt,
                 T* q_out, T* w_out, T& ceiling_out, unsigned* took = nullptr) {
  T b[3], c[3], L[3];
  rotate_t(q0, w, true, b);
  c[0] = b[0] * I[0]; c[1] = b[1] * I[1]; c[2] = b[2] * I[2];
  rotate_t(q0, c, false, L);
  T most = ceiling_in;
  const bool fresh = val(w[0]) != val(left[0]) || val(w[1]) != val(left[1]) || val(w[2]) != val(left[2]);
  if (fresh) most = w[0] * L[0] + w[1] * L[1] + w[2] * L[2];
  T q4[4], v[3], q[4];
  advance_t(q0, w, dt * 0.25, q4);
  angular_t(q4, inv, L, v);
  advance_t(q0, v, dt * 0.5, q);
  angular_t(q, inv, L, v);
  const bool over = val(v[0] * L[0] + v[1] * L[1] + v[2] * L[2]) > val(most) * (1 + ROUNDED);
  if (over) {
    T u[4];
    untwist_t(q, inv, L, most, u);
    for (int k = 0; k < 4; k++) q[k] = u[k];
    angular_t(q, inv, L, v);
  }
  // The energy check above compares within 2^-44 of the ceiling: a perturbation at rounding
  // level can flip it, and its outcomes differ by a turn of that size, so only whether the
  // ceiling was taken afresh is a branch of the record.
  if (took) *took = fresh ? 1u : 0u;
  for (int k = 0; k < 4; k++) q_out[k] = q[k];
  for (int k = 0; k < 3; k++) w_out[k] = v[k];
  ceiling_out = most;
}

// src/contact.js:1168-1182 (turned): the move over the step and the turn vector.
template <class T>
HD void turned_t(const T* p, const T* st, const T* q, const T* q0, T* out) {
  const T ax = q[0], ay = q[1], az = q[2], aw = q[3];
  const T bx = -q0[0], by = -q0[1], bz = -q0[2], bw = q0[3];
  T ux = aw * bx + ax * bw + ay * bz - az * by;
  T uy = aw * by + ay * bw + az * bx - ax * bz;
  T uz = aw * bz + az * bw + ax * by - ay * bx;
  T uw = aw * bw - ax * bx - ay * by - az * bz;
  if (val(uw) < 0) { ux = -ux; uy = -uy; uz = -uz; uw = -uw; }
  const T sine = sqrt_of(ux * ux + uy * uy + uz * uz);
  T k;
  if (val(sine) == 0) k = zero_of(sine);
  else k = 2 * atan2_of(sine, uw) / sine;
  out[0] = p[0] - st[0]; out[1] = p[1] - st[1]; out[2] = p[2] - st[2];
  out[3] = ux * k; out[4] = uy * k; out[5] = uz * k;
}

// ---------------------------------------------------------------------------------
// A held manifold placed from its bodies (src/contact.js:468-508, the REUSE rebuild),
// its tangent basis (src/contact.js:1283-1292). In: a's p 3 and q 4, b's p 3 and q 4.
// Out: normal 3, t1 3, t2 3, then each point's place 3 and separation 1.
struct Held {
  const double* ln;    // normal in a's frame
  const double* ends;  // per point: a's end in a's frame 3, b's end in b's frame 3
  double ra, rb;       // a's round radius toward b, b's (negated) toward a
  int count;
};
template <class T>
HD void place_t(const Held& h, const T* pa, const T* qa, const T* pb, const T* qb, T* out) {
  T A[9], B[9];
  axes_t(qa, A);
  axes_t(qb, B);
  const double lx = h.ln[0], ly = h.ln[1], lz = h.ln[2];
  const T nx = A[0] * lx + A[3] * ly + A[6] * lz;
  const T ny = A[1] * lx + A[4] * ly + A[7] * lz;
  const T nz = A[2] * lx + A[5] * ly + A[8] * lz;
  out[0] = nx; out[1] = ny; out[2] = nz;
  // basis: the branch is picked by the values, as the reference picks it
  const double fx = __builtin_fabs(val(nx)), fy = __builtin_fabs(val(ny)), fz = __builtin_fabs(val(nz));
  T tx, ty, tz;
  if (fx <= fy && fx <= fz) { tx = zero_of(nx); ty = nz; tz = -ny; }
  else if (fy <= fz) { tx = -nz; ty = zero_of(nx); tz = nx; }
  else { tx = ny; ty = -nx; tz = zero_of(nx); }
  const T l = 1 / sqrt_of(tx * tx + ty * ty + tz * tz);
  tx = tx * l; ty = ty * l; tz = tz * l;
  out[3] = tx; out[4] = ty; out[5] = tz;
  out[6] = ny * tz - nz * ty; out[7] = nz * tx - nx * tz; out[8] = nx * ty - ny * tx;
  for (int k = 0; k < h.count; k++) {
    const double* e = h.ends + 6 * k;
    const T pax = pa[0] + A[0] * e[0] + A[3] * e[1] + A[6] * e[2] + nx * h.ra;
    const T pay = pa[1] + A[1] * e[0] + A[4] * e[1] + A[7] * e[2] + ny * h.ra;
    const T paz = pa[2] + A[2] * e[0] + A[5] * e[1] + A[8] * e[2] + nz * h.ra;
    const T pbx = pb[0] + B[0] * e[3] + B[3] * e[4] + B[6] * e[5] + nx * h.rb;
    const T pby = pb[1] + B[1] * e[3] + B[4] * e[4] + B[7] * e[5] + ny * h.rb;
    const T pbz = pb[2] + B[2] * e[3] + B[5] * e[4] + B[8] * e[5] + nz * h.rb;
    const T s = (pbx - pax) * nx + (pby - pay) * ny + (pbz - paz) * nz, half = 0.5 * s;
    T* o = out + 9 + 4 * k;
    if (h.rb != 0) {
      o[0] = pbx - nx * half; o[1] = pby - ny * half; o[2] = pbz - nz * half;
    } else if (h.ra != 0) {
      o[0] = pax + nx * half; o[1] = pay + ny * half; o[2] = paz + nz * half;
    } else {
      o[0] = (pax + pbx) * 0.5; o[1] = (pay + pby) * 0.5; o[2] = (paz + pbz) * 0.5;
    }
    o[3] = s;
  }
}

// ---------------------------------------------------------------------------------
// Columns.

// Bodies, by global slot (a world's slots after its body offset).
struct Bodies {
  const uint8_t *type, *shape, *tumbles, *capped;  // tumbles: a lopsided free dynamic body; capped: spin capped
  const double *rounded, *gravity_scale, *damping, *angular_damping;
  const double *mass0, *inv_mass0, *inertia0, *inv_inertia0;  // as the world had them
  const double *mass, *friction, *restitution;                // the parameters
  double *inv_mass, *inertia, *inv_inertia;                   // from the mass
  double *p, *q, *v, *w, *left, *ceiling;                     // the state
  double *axes, *inv_i, *start, *start_q;                     // the step's
  uint64_t* branch;  // the forward pass's branches, a record a body (or nullptr: none kept)
};

// Rows: every solved manifold of every step, held (global row numbers).
struct Rows {
  const int *a, *b, *src;  // slots in the world; src: the row whose store this one carries, or -1
  const uint8_t* count;
  const uint32_t* pid;
  const double *local_normal, *ends, *tolerance, *patch;
  const double* shift;  // per point: where b's end was carried from at the moment of impact (0 if not met so)
  // what each row stored at its step's end (src/contact.js:1091-1115): the carried impulses
  double *store_impulse, *store_friction, *store_twist, *store_slip;
  uint64_t* branch;  // the forward pass's branches, a record a row (or nullptr: none kept)
};

// The branches a forward pass took (its active set: anchor weights, target cases, Murty sets,
// friction bias and cone, roll and spin clamps, bounces, spin caps, the tumble's branches),
// folded into one record a row and one a body in the order they are taken. Each row and each
// body is worked on by one thread at a time, so the records need no atomics; two passes took
// the same branches when every record is equal.
HD void note(uint64_t* record, uint64_t code) {
  if (record) *record = (*record ^ (code + 0x9e3779b97f4a7c15ull)) * 1099511628211ull;
}

// The current step's rows (by working index: the row less the step's first row).
// The forward pass's working columns (per point: 4 slots a row).
struct Work {
  double* base;  // the columns, one after another, each n rows wide
  size_t n;
  uint8_t* sliding;
  HD double* point() const { return base + 0 * n; }  // 12 a row
  HD double* separation() const { return base + 12 * n; }  // 4 a row
  HD double* normal() const { return base + 16 * n; }  // 3 a row
  HD double* t1() const { return base + 19 * n; }  // 3 a row
  HD double* t2() const { return base + 22 * n; }  // 3 a row
  HD double* mu() const { return base + 25 * n; }  // 1 a row
  HD double* e() const { return base + 26 * n; }  // 1 a row
  HD double* impulse() const { return base + 27 * n; }  // 4 a row
  HD double* tangent() const { return base + 31 * n; }  // 2 a row
  HD double* roll() const { return base + 33 * n; }  // 2 a row
  HD double* spin() const { return base + 35 * n; }  // 1 a row
  HD double* drift() const { return base + 36 * n; }  // 2 a row
  HD double* most() const { return base + 38 * n; }  // 4 a row
  HD double* relvel() const { return base + 42 * n; }  // 4 a row
  HD double* now() const { return base + 46 * n; }  // 4 a row
  HD double* ra() const { return base + 50 * n; }  // 12 a row
  HD double* rb() const { return base + 62 * n; }  // 12 a row
  HD double* la() const { return base + 74 * n; }  // 12 a row
  HD double* lb() const { return base + 86 * n; }  // 12 a row
  HD double* adjusted() const { return base + 98 * n; }  // 4 a row
  HD double* nmass() const { return base + 102 * n; }  // 4 a row
  HD double* jac() const { return base + 106 * n; }  // 48 a row
  HD double* block() const { return base + 154 * n; }  // 16 a row
  HD double* lever() const { return base + 170 * n; }  // 4 a row
  HD double* anchor_a() const { return base + 174 * n; }  // 3 a row
  HD double* anchor_b() const { return base + 177 * n; }  // 3 a row
  HD double* fjac() const { return base + 180 * n; }  // 24 a row
  HD double* tm() const { return base + 204 * n; }  // 3 a row
  HD double* rm() const { return base + 207 * n; }  // 2 a row
  HD double* sm() const { return base + 209 * n; }  // 1 a row
  HD double* bounce() const { return base + 210 * n; }  // 56 a row
  static constexpr int WIDTH = 266;
};
constexpr int BOUNCE = 14;

// A world's step: its bodies from boff, rows [r0, r1) (working index from w0), its
// colour runs and the order narrow() found the rows in.
struct StepView {
  int boff, nbodies, r0, r1, w0, substeps;
  double dt, gx, gy, gz;
  const int* starts;  // STARTS, relative to r0
  const int* narrow;  // [r1 - r0]: the rows (relative) in narrow order
};

struct Ctx {
  int tid, threads;
  // This thread's block of ROW_SLOT doubles for a row's adjoints while its reverse runs (the
  // workgroup's LDS on the device); only the backward pass uses it.
  double* slot;
};


// ---------------------------------------------------------------------------------
// The held step, forward. Each function follows the soft step's (soft.hip), which
// follows the reference's lines cited there, operation for operation.

struct Sizes {
  const double* segment;  // a capsule's half segment (src/contact.js:1213), else 0
};

// A held row's places, its mixed friction and restitution (src/contact.js:889-921 emit),
// then the impulses its pair carried (src/contact.js:999-1036 carry).
HD Held held_of(const Bodies& B, const Rows& R, int boff, int r) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  return Held{R.local_normal + 3 * r, R.ends + 6 * P * r, B.rounded[ga], -B.rounded[gb], R.count[r]};
}

HD void place_row(const Bodies& B, const Rows& R, const Work& W, int boff, int r, int i) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const Held h = held_of(B, R, boff, r);
  double out[9 + 4 * P];
  place_t<double>(h, B.p + 3 * ga, B.q + 4 * ga, B.p + 3 * gb, B.q + 4 * gb, out);
  for (int k = 0; k < 3; k++) {
    W.normal()[3 * i + k] = out[k]; W.t1()[3 * i + k] = out[3 + k]; W.t2()[3 * i + k] = out[6 + k];
  }
  for (int k = 0; k < h.count; k++) {
    const int j = P * i + k;
    for (int c = 0; c < 3; c++) W.point()[3 * j + c] = out[9 + 4 * k + c];
    W.separation()[j] = out[9 + 4 * k + 3];
  }
  W.mu()[i] = dsqrt(B.friction[ga] * B.friction[gb]);
  W.e()[i] = js_max(B.restitution[ga], B.restitution[gb]);
}

HD void carry_row(const Rows& R, const Work& W, int r, int i) {
  const int count = R.count[r];
  W.roll()[2 * i] = 0; W.roll()[2 * i + 1] = 0; W.spin()[i] = 0;
  W.tangent()[2 * i] = 0; W.tangent()[2 * i + 1] = 0; W.drift()[2 * i] = 0; W.drift()[2 * i + 1] = 0;
  for (int k = 0; k < count; k++) W.impulse()[P * i + k] = 0;
  const int s = R.src[r];
  if (s < 0) return;
  const double *t1 = W.t1() + 3 * i, *t2 = W.t2() + 3 * i, *n = W.normal() + 3 * i;
  const double* L = R.store_twist + 3 * s;
  if (count <= 2) {
    W.roll()[2 * i] = L[0] * t1[0] + L[1] * t1[1] + L[2] * t1[2];
    W.roll()[2 * i + 1] = L[0] * t2[0] + L[1] * t2[1] + L[2] * t2[2];
  }
  W.spin()[i] = L[0] * n[0] + L[1] * n[1] + L[2] * n[2];
  const double* F = R.store_friction + 3 * s;
  W.tangent()[2 * i] = F[0] * t1[0] + F[1] * t1[1] + F[2] * t1[2];
  W.tangent()[2 * i + 1] = F[0] * t2[0] + F[1] * t2[1] + F[2] * t2[2];
  bool same = false;
  for (int k = 0; k < count; k++) {
    const uint32_t id = R.pid[P * r + k];
    for (int j = 0; j < R.count[s]; j++) {
      if (R.pid[P * s + j] != id) continue;
      W.impulse()[P * i + k] = R.store_impulse[P * s + j];
      same = true;
      break;
    }
  }
  if (same) {
    const double* S = R.store_slip + 3 * s;
    W.drift()[2 * i] = S[0] * t1[0] + S[1] * t1[1] + S[2] * t1[2];
    W.drift()[2 * i + 1] = S[0] * t2[0] + S[1] * t2[1] + S[2] * t2[2];
  }
}

// src/contact.js:1499-1515 (inertias) for a dynamic body: the axes from q, then
// R diag(I^-1) R^T.
template <class T>
HD void world_inverse_t(const T* q, const T* inv, T* out) {
  T ax[9];
  axes_t(q, ax);
  const T x0 = ax[0], x1 = ax[1], x2 = ax[2], y0 = ax[3], y1 = ax[4], y2 = ax[5], z0 = ax[6], z1 = ax[7], z2 = ax[8];
  out[0] = x0 * x0 * inv[0] + y0 * y0 * inv[1] + z0 * z0 * inv[2];
  out[1] = x0 * x1 * inv[0] + y0 * y1 * inv[1] + z0 * z1 * inv[2];
  out[2] = x0 * x2 * inv[0] + y0 * y2 * inv[1] + z0 * z2 * inv[2];
  out[3] = x1 * x1 * inv[0] + y1 * y1 * inv[1] + z1 * z1 * inv[2];
  out[4] = x1 * x2 * inv[0] + y1 * y2 * inv[1] + z1 * z2 * inv[2];
  out[5] = x2 * x2 * inv[0] + y2 * y2 * inv[1] + z2 * z2 * inv[2];
}

HD void inertia_body(const Bodies& B, int g) {
  double* out = B.inv_i + 6 * g;
  if (B.type[g] != DYNAMIC) {
    for (int k = 0; k < 6; k++) out[k] = 0;
    return;
  }
  world_inverse_t<double>(B.q + 4 * g, B.inv_inertia + 3 * g, out);
}

// src/contact.js:1709-1722 (turning, along)
HD double along(const double* I, double x, double y, double z) {
  return I[0] * x * x + I[3] * y * y + I[5] * z * z + 2 * (I[1] * x * y + I[2] * x * z + I[4] * y * z);
}
HD double turning(const double* Ia, const double* Ib, double x, double y, double z) {
  const double k = along(Ia, x, y, z) + along(Ib, x, y, z);
  return k > 0 ? 1 / k : 0;
}

// src/contact.js:1565-1581 (lever)
HD double lever(double* jac, const double* Ia, const double* Ib, double ima, double imb, double rax, double ray,
                double raz, double rbx, double rby, double rbz, double dx, double dy, double dz) {
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

// src/contact.js:1199-1220 (core): the anchor in the body's frame into out; returns its
// world offset at the step's start.
HD V3 core(const Bodies& B, const Sizes& Z, int g, double x, double y, double z, double* out) {
  const int shape = B.shape[g];
  if (shape == SPHERE_SHAPE) {
    out[0] = 0; out[1] = 0; out[2] = 0;
    return V3{0, 0, 0};
  }
  const double* q = B.q + 4 * g;
  V3 v = rotate(q, x, y, z, true);
  if (shape != CAPSULE_SHAPE) {
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
    return V3{x, y, z};
  }
  const double half = Z.segment[g];
  const double c = v.y < -half ? -half : v.y > half ? half : v.y;
  out[0] = 0; out[1] = c; out[2] = 0;
  return rotate(q, 0, c, 0, false);
}

// src/contact.js:1346-1445 (prepareRows) for row i; the bodies are at the step's start.
HD void prepare_row(const Bodies& B, const Sizes& Z, const Work& W, const Rows& R, int boff, int r, int i) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const double* Ia = B.inv_i + 6 * ga;
  const double* Ib = B.inv_i + 6 * gb;
  const double *pa = B.p + 3 * ga, *pb = B.p + 3 * gb;
  const double *va = B.v + 3 * ga, *vb = B.v + 3 * gb;
  const double *wa = B.w + 3 * ga, *wb = B.w + 3 * gb;
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  const int i3 = 3 * i;
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const int count = R.count[r];
  if (count <= 2) {
    W.rm()[2 * i] = turning(Ia, Ib, W.t1()[i3], W.t1()[i3 + 1], W.t1()[i3 + 2]);
    W.rm()[2 * i + 1] = turning(Ia, Ib, W.t2()[i3], W.t2()[i3 + 1], W.t2()[i3 + 2]);
  }
  W.sm()[i] = turning(Ia, Ib, nx, ny, nz);
  const double tol = R.tolerance[r];
  double weight = 0, ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  double ex = 0, ey = 0, ez = 0, fx = 0, fy = 0, fz = 0;
  for (int k = 0; k < count; k++) {
    const int j = P * i + k, j3 = 3 * j;
    const double px = W.point()[j3], py = W.point()[j3 + 1], pz = W.point()[j3 + 2];
    const double rax = px - pa[0], ray = py - pa[1], raz = pz - pa[2];
    const double* sh = R.shift + 3 * (P * r + k);
    const double rbx = px + sh[0] - pb[0], rby = py + sh[1] - pb[1], rbz = pz + sh[2] - pb[2];
    W.ra()[j3] = rax; W.ra()[j3 + 1] = ray; W.ra()[j3 + 2] = raz;
    W.rb()[j3] = rbx; W.rb()[j3 + 1] = rby; W.rb()[j3 + 2] = rbz;
    const V3 ca = core(B, Z, ga, rax, ray, raz, W.la() + j3);
    const V3 cb = core(B, Z, gb, rbx, rby, rbz, W.lb() + j3);
    W.adjusted()[j] = W.separation()[j] - ((cb.x - ca.x) * nx + (cb.y - ca.y) * ny + (cb.z - ca.z) * nz);
    W.nmass()[j] = lever(W.jac() + J * j, Ia, Ib, ima, imb, rax, ray, raz, rbx, rby, rbz, nx, ny, nz);
    W.most()[j] = 0;
    const double s = W.separation()[j] / tol, near = s <= 1 ? 1 : s >= 2 ? 0 : 2 - s;
    note(R.branch ? R.branch + r : nullptr, 0x100u | (s <= 1 ? 0u : s >= 2 ? 1u : 2u));
    weight += near;
    ax += near * rax; ay += near * ray; az += near * raz; bx += near * rbx; by += near * rby; bz += near * rbz;
    ex += rax; ey += ray; ez += raz; fx += rbx; fy += rby; fz += rbz;
    const double dvx = vb[0] + wb[1] * rbz - wb[2] * rby - va[0] - wa[1] * raz + wa[2] * ray;
    const double dvy = vb[1] + wb[2] * rbx - wb[0] * rbz - va[1] - wa[2] * rax + wa[0] * raz;
    const double dvz = vb[2] + wb[0] * rby - wb[1] * rbx - va[2] - wa[0] * ray + wa[1] * rax;
    W.relvel()[j] = dvx * nx + dvy * ny + dvz * nz;
  }
  if (count >= 2) {
    const double* jac = W.jac() + J * (P * i);
    double* K = W.block() + P * P * i;
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
  note(R.branch ? R.branch + r : nullptr, 0x200u | (weight == 0 ? 1u : 0u));
  if (weight == 0) { weight = count; ax = ex; ay = ey; az = ez; bx = fx; by = fy; bz = fz; }
  ax /= weight; ay /= weight; az /= weight; bx /= weight; by /= weight; bz /= weight;
  W.anchor_a()[i3] = ax; W.anchor_a()[i3 + 1] = ay; W.anchor_a()[i3 + 2] = az;
  W.anchor_b()[i3] = bx; W.anchor_b()[i3 + 1] = by; W.anchor_b()[i3 + 2] = bz;
  for (int k = 0; k < count; k++) {
    const int j3 = 3 * (P * i + k);
    double dx = W.ra()[j3] - ax, dy = W.ra()[j3 + 1] - ay, dz = W.ra()[j3 + 2] - az;
    const double al = dx * nx + dy * ny + dz * nz;
    dx -= al * nx; dy -= al * ny; dz -= al * nz;
    W.lever()[P * i + k] = dsqrt(dx * dx + dy * dy + dz * dz);
  }
  double* f = W.fjac() + FJ * i;
  lever(f, Ia, Ib, ima, imb, ax, ay, az, bx, by, bz, W.t1()[i3], W.t1()[i3 + 1], W.t1()[i3 + 2]);
  lever(f + 12, Ia, Ib, ima, imb, ax, ay, az, bx, by, bz, W.t2()[i3], W.t2()[i3 + 1], W.t2()[i3 + 2]);
  const double k11 = ima + imb + f[0] * f[6] + f[1] * f[7] + f[2] * f[8] + f[3] * f[9] + f[4] * f[10] + f[5] * f[11];
  const double k22 = ima + imb + f[12] * f[18] + f[13] * f[19] + f[14] * f[20] + f[15] * f[21] + f[16] * f[22]
      + f[17] * f[23];
  const double k12 = f[12] * f[6] + f[13] * f[7] + f[14] * f[8] + f[15] * f[9] + f[16] * f[10] + f[17] * f[11];
  const double det = k11 * k22 - k12 * k12;
  double* tm = W.tm() + i3;
  note(R.branch ? R.branch + r : nullptr, 0x300u | (det > 0 ? 1u : 0u));
  if (det > 0) {
    tm[0] = k22 / det; tm[1] = -k12 / det; tm[2] = k11 / det;
  } else {
    tm[0] = 0; tm[1] = 0; tm[2] = 0;
  }
}

// The two bodies' velocities for a row.
struct Pair {
  double v[12];  // va 3, wa 3, vb 3, wb 3
};
HD Pair gather(const Bodies& B, int ga, int gb) {
  Pair s;
  for (int k = 0; k < 3; k++) {
    s.v[k] = B.v[3 * ga + k]; s.v[3 + k] = B.w[3 * ga + k];
    s.v[6 + k] = B.v[3 * gb + k]; s.v[9 + k] = B.w[3 * gb + k];
  }
  return s;
}
// src/contact.js:2096-2104 (writeBack): a body that cannot move is never written.
HD void scatter(const Bodies& B, int ga, int gb, double ima, double imb, const Pair& s) {
  if (ima != 0)
    for (int k = 0; k < 3; k++) { B.v[3 * ga + k] = s.v[k]; B.w[3 * ga + k] = s.v[3 + k]; }
  if (imb != 0)
    for (int k = 0; k < 3; k++) { B.v[3 * gb + k] = s.v[6 + k]; B.w[3 * gb + k] = s.v[9 + k]; }
}

// s += (lin on v, I-weighted on w) for an impulse along n with turning terms jac[6..11].
HD void push_normal(Pair& s, double ima, double imb, double nx, double ny, double nz, const double* jac, double l) {
  const double px = nx * l, py = ny * l, pz = nz * l;
  s.v[0] -= ima * px; s.v[1] -= ima * py; s.v[2] -= ima * pz;
  s.v[6] += imb * px; s.v[7] += imb * py; s.v[8] += imb * pz;
  s.v[3] -= jac[6] * l; s.v[4] -= jac[7] * l; s.v[5] -= jac[8] * l;
  s.v[9] += jac[9] * l; s.v[10] += jac[10] * l; s.v[11] += jac[11] * l;
}
// wa -= Ia L, wb += Ib L
HD void push_turn(Pair& s, const double* Ia, const double* Ib, double lx, double ly, double lz) {
  s.v[3] -= Ia[0] * lx + Ia[1] * ly + Ia[2] * lz;
  s.v[4] -= Ia[1] * lx + Ia[3] * ly + Ia[4] * lz;
  s.v[5] -= Ia[2] * lx + Ia[4] * ly + Ia[5] * lz;
  s.v[9] += Ib[0] * lx + Ib[1] * ly + Ib[2] * lz;
  s.v[10] += Ib[1] * lx + Ib[3] * ly + Ib[4] * lz;
  s.v[11] += Ib[2] * lx + Ib[4] * ly + Ib[5] * lz;
}
HD double closing(const Pair& s, double nx, double ny, double nz, const double* jac) {
  return (s.v[6] - s.v[0]) * nx + (s.v[7] - s.v[1]) * ny + (s.v[8] - s.v[2]) * nz + s.v[9] * jac[3]
      + s.v[10] * jac[4] + s.v[11] * jac[5] - s.v[3] * jac[0] - s.v[4] * jac[1] - s.v[5] * jac[2];
}

// src/contact.js:1616-1663 (warmRows)
HD void warm_row(const Bodies& B, const Work& W, const Rows& R, int boff, int r, int i) {
  const int ga = boff + R.a[r], gb = boff + R.b[r], count = R.count[r];
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  const int i3 = 3 * i;
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const double t1x = W.t1()[i3], t1y = W.t1()[i3 + 1], t1z = W.t1()[i3 + 2];
  const double t2x = W.t2()[i3], t2y = W.t2()[i3 + 1], t2z = W.t2()[i3 + 2];
  Pair s = gather(B, ga, gb);
  for (int k = 0; k < count; k++) {
    const int j = P * i + k;
    push_normal(s, ima, imb, nx, ny, nz, W.jac() + J * j, W.impulse()[j]);
  }
  {
    const double* fj = W.fjac() + FJ * i;
    const double l1 = W.tangent()[2 * i], l2 = W.tangent()[2 * i + 1];
    const double px = t1x * l1 + t2x * l2, py = t1y * l1 + t2y * l2, pz = t1z * l1 + t2z * l2;
    s.v[0] -= ima * px; s.v[1] -= ima * py; s.v[2] -= ima * pz;
    s.v[6] += imb * px; s.v[7] += imb * py; s.v[8] += imb * pz;
    s.v[3] -= fj[6] * l1 + fj[18] * l2; s.v[4] -= fj[7] * l1 + fj[19] * l2; s.v[5] -= fj[8] * l1 + fj[20] * l2;
    s.v[9] += fj[9] * l1 + fj[21] * l2; s.v[10] += fj[10] * l1 + fj[22] * l2; s.v[11] += fj[11] * l1 + fj[23] * l2;
  }
  {
    const double r1 = W.roll()[2 * i], r2 = W.roll()[2 * i + 1], sp = W.spin()[i];
    const double lx = t1x * r1 + t2x * r2 + nx * sp, ly = t1y * r1 + t2y * r2 + ny * sp, lz = t1z * r1 + t2z * r2 + nz * sp;
    push_turn(s, B.inv_i + 6 * ga, B.inv_i + 6 * gb, lx, ly, lz);
  }
  scatter(B, ga, gb, ima, imb, s);
}

struct Soft {
  double bias_rate, mass_scale, impulse_scale;
};
// src/contact.js:1590-1597
HD Soft make_soft(double hertz, double zeta, double h) {
  const double omega = 2 * MATH_PI * hertz;
  const double a1 = 2 * zeta + h * omega, a2 = h * omega * a1, a3 = 1 / (1 + a2);
  return Soft{omega / a1, a2 * a3, a3};
}


// src/contact.js:1672-1724 (pushes): Gaussian elimination without pivoting on the rows
// and columns of `set`, in rising order; the rest 0.
template <int N>
HD void pushes(const double (&K)[N][N], int set, const double (&give)[N], const double (&wanted)[N],
               double (&solved)[N]) {
  double A[N][N], Rr[N], X[N];
  for (int r = 0; r < N; r++) {
    for (int c = 0; c < N; c++) A[r][c] = K[r][c];
    A[r][r] *= 1 + give[r];
    Rr[r] = wanted[r];
    solved[r] = 0;
    X[r] = 0;
  }
  for (int col = 0; col < N; col++) {
    if (!((set >> col) & 1)) continue;
    const double pivot = A[col][col];
    for (int r = col + 1; r < N; r++) {
      if (!((set >> r) & 1)) continue;
      const double f = A[r][col] / pivot;
      for (int c = col + 1; c < N; c++)
        if ((set >> c) & 1) A[r][c] -= f * A[col][c];
      Rr[r] -= f * Rr[col];
    }
  }
  for (int r = N - 1; r >= 0; r--) {
    if (!((set >> r) & 1)) continue;
    double x = Rr[r];
    for (int c = r + 1; c < N; c++)
      if ((set >> c) & 1) x -= A[r][c] * X[c];
    X[r] = x / A[r][r];
    solved[r] = X[r];
  }
}

// What a row's solve or relax read: its bodies' velocities and its impulses as the sweep
// found them (the tape), and its points' separations now.
template <int N>
struct RowIn {
  Pair s;
  double impulse[N], o1, o2, r1, r2, s0, most[N], sep[N];
};

// Everything the row worked out, kept for its reverse.
template <int N>
struct RowFwd {
  double target[N], mscale[N], iscale[N], give[N];
  int tcase[N];  // 0: s > 0; 1: bias at the rate; 2: bias at -max_bias; 3: no bias
  bool single;
  double closing[N], raw[N], solved[N], next[N], dd[N];
  int set;
  uint32_t first, larger;
  double pressed, twisted;
  Pair mid;  // velocities after the normal pushes
  int fbias;  // 0 none, 1 at the rate, 2 capped
  double b1, b2, bb, v1, v2, n1, n2, limit, al, x, y, ll, t1_out, t2_out;
  bool slide;
  double u[3], rn1, rn2, roll_most, r1_out, r2_out;
  bool roll_capped;
  double dw[3], slimit, sraw, s_out;
  int sclamp;  // 0 none, 1 at +limit, -1 at -limit
  double least;
};

// src/contact.js:1805-2090 (solveRows) for one row of N points: the forward, from `in`,
// leaving `out` (the velocities after) and the row's results in f.
template <int N>
HD void solve_math(const Work& W, const Rows& R, int r, int i, double ima, double imb, const double* Ia,
                   const double* Ib, const Soft& soft, double h, bool bias, const RowIn<N>& in, RowFwd<N>& f,
                   Pair& s) {
  const double inv = 1 / h;
  const double springy = soft.impulse_scale / soft.mass_scale + COMPLIANCE;
  const double tol = R.tolerance[r], mu = W.mu()[i], patch = R.patch[r];
  const int i3 = 3 * i;
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const double max_bias = tol * inv;
  double least = __builtin_inf();
  for (int k = 0; k < N; k++) {
    const double sk = in.sep[k];
    if (sk < least) least = sk;
    if (sk > 0) {
      f.target[k] = sk * inv; f.mscale[k] = 1; f.iscale[k] = 0; f.give[k] = COMPLIANCE; f.tcase[k] = 0;
    } else if (bias) {
      const double t = soft.bias_rate * sk;
      f.target[k] = js_max(t, -max_bias);
      f.tcase[k] = t > -max_bias ? 1 : (t == -max_bias ? 1 : 2);
      f.mscale[k] = soft.mass_scale; f.iscale[k] = soft.impulse_scale; f.give[k] = springy;
    } else {
      f.target[k] = 0; f.mscale[k] = 1; f.iscale[k] = 0; f.give[k] = COMPLIANCE; f.tcase[k] = 3;
    }
  }
  f.least = least;
  s = in.s;
  const double* Kg = W.block() + P * P * i;
  f.single = N == 1 || !(Kg[0] > 0);
  f.set = 0;
  if (f.single) {
    for (int k = 0; k < N; k++) {
      const double* jac = W.jac() + J * (P * i + k);
      const double vn = closing(s, nx, ny, nz, jac);
      const double old = in.impulse[k];
      double next = old - W.nmass()[P * i + k] * f.mscale[k] * (vn + f.target[k]) - f.iscale[k] * old;
      f.raw[k] = next;
      if (next < 0) next = 0;
      f.closing[k] = vn; f.solved[k] = next;
    }
  } else {
    double K[N][N], wanted[N];
    for (int k = 0; k < N; k++)
      for (int l = 0; l < N; l++) K[k][l] = Kg[P * k + l];
    for (int k = 0; k < N; k++) f.closing[k] = closing(s, nx, ny, nz, W.jac() + J * (P * i + k));
    int set = 0;
    for (int k = 0; k < N; k++) {
      double q = -f.closing[k] - f.target[k];
      for (int l = 0; l < N; l++) q += K[k][l] * in.impulse[l];
      wanted[k] = q;
      if (q > 0) set |= 1 << k;
    }
    for (int round = 0;; round++) {
      pushes<N>(K, set, f.give, wanted, f.solved);
      int flip = -1;
      for (int k = 0; k < N; k++) {
        if (flip >= 0) continue;
        if ((set >> k) & 1) {
          if (f.solved[k] < 0) flip = k;
        } else {
          double q = -wanted[k];
          for (int l = 0; l < N; l++) q += K[k][l] * f.solved[l];
          if (q < 0) flip = k;
        }
      }
      if (flip < 0 || round == (1 << N)) break;
      set ^= 1 << flip;
    }
    f.set = set;
  }
  f.first = 0; f.larger = 0;
  for (int k = 0; k < N; k++) {
    const double old = in.impulse[k], next = f.solved[k] > 0 ? f.solved[k] : 0;
    const double dd = next - old;
    f.next[k] = next; f.dd[k] = dd;
    if (dd > 0 && in.most[k] == 0) f.first |= 1u << k;
    if (dd > in.most[k]) f.larger |= 1u << k;
    push_normal(s, ima, imb, nx, ny, nz, W.jac() + J * (P * i + k), dd);
  }
  f.mid = s;
  const double t1x = W.t1()[i3], t1y = W.t1()[i3 + 1], t1z = W.t1()[i3 + 2];
  const double t2x = W.t2()[i3], t2y = W.t2()[i3 + 1], t2z = W.t2()[i3 + 2];
  double pressed = 0, twisted = 0;
  for (int k = 0; k < N; k++) {
    pressed += f.next[k];
    twisted += W.lever()[P * i + k] * f.next[k];
  }
  f.pressed = pressed; f.twisted = twisted;
  {
    const double* fj = W.fjac() + FJ * i;
    const double* g = fj + 12;
    const double dvx = s.v[6] - s.v[0], dvy = s.v[7] - s.v[1], dvz = s.v[8] - s.v[2];
    double v1 = dvx * t1x + dvy * t1y + dvz * t1z + s.v[9] * fj[3] + s.v[10] * fj[4] + s.v[11] * fj[5]
        - s.v[3] * fj[0] - s.v[4] * fj[1] - s.v[5] * fj[2];
    double v2 = dvx * t2x + dvy * t2y + dvz * t2z + s.v[9] * g[3] + s.v[10] * g[4] + s.v[11] * g[5]
        - s.v[3] * g[0] - s.v[4] * g[1] - s.v[5] * g[2];
    const double s1 = W.drift()[2 * i], s2 = W.drift()[2 * i + 1];
    f.fbias = 0;
    if (bias && (s1 != 0 || s2 != 0)) {
      double b1 = soft.bias_rate * s1, b2 = soft.bias_rate * s2;
      const double bb = b1 * b1 + b2 * b2;
      f.b1 = b1; f.b2 = b2; f.bb = bb;
      f.fbias = 1;
      if (bb > max_bias * max_bias) {
        const double scale = max_bias / dsqrt(bb);
        b1 *= scale; b2 *= scale;
        f.fbias = 2;
      }
      v1 += b1; v2 += b2;
    }
    f.v1 = v1; f.v2 = v2;
    const double tm0 = W.tm()[i3], tm1 = W.tm()[i3 + 1], tm2 = W.tm()[i3 + 2];
    double n1 = in.o1 - (tm0 * v1 + tm1 * v2), n2 = in.o2 - (tm1 * v1 + tm2 * v2);
    f.n1 = n1; f.n2 = n2;
    const double limit = mu * pressed, l2 = n1 * n1 + n2 * n2;
    f.limit = limit;
    f.slide = false;
    f.al = 0; f.x = n1; f.y = n2; f.ll = l2;
    if (l2 > limit * limit) {
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
      f.al = al; f.x = x; f.y = y; f.ll = ll;
      const double scale = limit / dsqrt(ll);
      n1 = x * scale; n2 = y * scale;
      f.slide = true;
    }
    f.t1_out = n1; f.t2_out = n2;
    const double d1 = n1 - in.o1, d2 = n2 - in.o2;
    const double px = t1x * d1 + t2x * d2, py = t1y * d1 + t2y * d2, pz = t1z * d1 + t2z * d2;
    s.v[0] -= ima * px; s.v[1] -= ima * py; s.v[2] -= ima * pz;
    s.v[6] += imb * px; s.v[7] += imb * py; s.v[8] += imb * pz;
    s.v[3] -= fj[6] * d1 + g[6] * d2; s.v[4] -= fj[7] * d1 + g[7] * d2; s.v[5] -= fj[8] * d1 + g[8] * d2;
    s.v[9] += fj[9] * d1 + g[9] * d2; s.v[10] += fj[10] * d1 + g[10] * d2; s.v[11] += fj[11] * d1 + g[11] * d2;
  }
  f.r1_out = 0; f.r2_out = 0; f.roll_capped = false;
  if (N <= 2) {
    const double most = patch * pressed;
    f.roll_most = most;
    const double ux = s.v[9] - s.v[3], uy = s.v[10] - s.v[4], uz = s.v[11] - s.v[5];
    f.u[0] = ux; f.u[1] = uy; f.u[2] = uz;
    double n1 = in.r1 - W.rm()[2 * i] * (ux * t1x + uy * t1y + uz * t1z);
    double n2 = in.r2 - W.rm()[2 * i + 1] * (ux * t2x + uy * t2y + uz * t2z);
    f.rn1 = n1; f.rn2 = n2;
    const double l2 = n1 * n1 + n2 * n2;
    if (l2 > most * most) {
      const double scale = most / dsqrt(l2);
      n1 *= scale; n2 *= scale;
      f.roll_capped = true;
    }
    f.r1_out = n1; f.r2_out = n2;
    const double d1 = n1 - in.r1, d2 = n2 - in.r2;
    push_turn(s, Ia, Ib, t1x * d1 + t2x * d2, t1y * d1 + t2y * d2, t1z * d1 + t2z * d2);
  }
  {
    const double limit = N == 1 ? SPIN_DISC * mu * patch * pressed : mu * twisted;
    f.slimit = limit;
    const double s0 = in.s0;
    const double dwx = s.v[9] - s.v[3], dwy = s.v[10] - s.v[4], dwz = s.v[11] - s.v[5];
    f.dw[0] = dwx; f.dw[1] = dwy; f.dw[2] = dwz;
    double s1 = s0 - W.sm()[i] * (dwx * nx + dwy * ny + dwz * nz);
    f.sraw = s1;
    f.sclamp = 0;
    if (s1 > limit) { s1 = limit; f.sclamp = 1; }
    else if (s1 < -limit) { s1 = -limit; f.sclamp = -1; }
    f.s_out = s1;
    const double ds = s1 - s0;
    push_turn(s, Ia, Ib, nx * ds, ny * ds, nz * ds);
  }
}


// The separations a row's points have now (src/contact.js:1838-1862): a known sweep reads
// the ones the last relax left; otherwise each point's ends, carried with their bodies
// since the step began, measured along the normal.
template <int N>
HD void separations(const Bodies& B, const Work& W, int ga, int gb, int i, bool known, double (&sep)[N]) {
  if (known) {
    for (int k = 0; k < N; k++) sep[k] = W.now()[P * i + k];
    return;
  }
  const int i3 = 3 * i;
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const double *pa = B.p + 3 * ga, *pb = B.p + 3 * gb;
  const double *qa = B.start + 3 * ga, *qb = B.start + 3 * gb;
  const double dx = (pb[0] - qb[0]) - (pa[0] - qa[0]);
  const double dy = (pb[1] - qb[1]) - (pa[1] - qa[1]);
  const double dz = (pb[2] - qb[2]) - (pa[2] - qa[2]);
  const double dn = dx * nx + dy * ny + dz * nz;
  const V3 na = rotate(B.q + 4 * ga, nx, ny, nz, true);
  const V3 nb = rotate(B.q + 4 * gb, nx, ny, nz, true);
  for (int k = 0; k < N; k++) {
    const int j3 = 3 * (P * i + k);
    sep[k] = dn + (W.lb()[j3] * nb.x + W.lb()[j3 + 1] * nb.y + W.lb()[j3 + 2] * nb.z)
        - (W.la()[j3] * na.x + W.la()[j3 + 1] * na.y + W.la()[j3 + 2] * na.z) + W.adjusted()[P * i + k];
  }
}

// A solve (bias) or relax sweep's row, forward; its inputs onto `tape` if given.
template <int N>
HD void row_forward(const Bodies& B, const Work& W, const Rows& R, int boff, int r, int i, const Soft& soft,
                    double h, bool bias, bool known, const TapeRow& tape) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  RowIn<N> in;
  in.s = gather(B, ga, gb);
  for (int k = 0; k < N; k++) {
    in.impulse[k] = W.impulse()[P * i + k];
    in.most[k] = W.most()[P * i + k];
  }
  in.o1 = W.tangent()[2 * i]; in.o2 = W.tangent()[2 * i + 1];
  in.r1 = N <= 2 ? W.roll()[2 * i] : 0; in.r2 = N <= 2 ? W.roll()[2 * i + 1] : 0;
  in.s0 = W.spin()[i];
  separations<N>(B, W, ga, gb, i, known, in.sep);
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  RowFwd<N> f;
  Pair s;
  solve_math<N>(W, R, r, i, ima, imb, B.inv_i + 6 * ga, B.inv_i + 6 * gb, soft, h, bias, in, f, s);
  scatter(B, ga, gb, ima, imb, s);
  if (R.branch) {
    uint64_t code = (f.single ? 1u : 0u) | (f.slide ? 2u : 0u) | (f.roll_capped ? 4u : 0u)
        | ((uint64_t)(f.sclamp == 1 ? 1 : f.sclamp == -1 ? 2 : 0) << 3) | ((uint64_t)f.fbias << 5)
        | ((uint64_t)f.set << 8) | ((uint64_t)f.first << 12)
        | ((uint64_t)(bias ? 1 : 0) << 40) | ((uint64_t)(known ? 1 : 0) << 41) | ((uint64_t)N << 42);
    for (int k = 0; k < N; k++) code |= (uint64_t)f.tcase[k] << (20 + 2 * k);
    note(R.branch + r, code);
  }
  if (tape.p) {
    const TapeRow& T = tape;
    for (int k = 0; k < 12; k++) { T[T_S + k] = in.s.v[k]; T[T_MID + k] = f.mid.v[k]; }
    uint32_t flags = (f.single ? 1u : 0u) | (f.slide ? 2u : 0u) | (f.roll_capped ? 4u : 0u)
        | ((uint32_t)(f.sclamp == 1 ? 1 : f.sclamp == -1 ? 2 : 0) << 3) | ((uint32_t)f.fbias << 5)
        | ((uint32_t)f.set << 8) | ((uint32_t)f.first << 12);
    for (int k = 0; k < N; k++) {
      T[T_IMP + k] = in.impulse[k]; T[T_MOST + k] = in.most[k]; T[T_TARGET + k] = f.target[k];
      T[T_CLOSING + k] = f.closing[k]; T[T_RAW + k] = f.single ? f.raw[k] : 0; T[T_SOLVED + k] = f.solved[k];
      T[T_NEXT + k] = f.next[k]; T[T_DD + k] = f.dd[k];
      flags |= (uint32_t)f.tcase[k] << (16 + 2 * k);
    }
    T[T_O1] = in.o1; T[T_O2] = in.o2; T[T_R1] = in.r1; T[T_R2] = in.r2; T[T_S0] = in.s0;
    T[T_FLAGS] = flags; T[T_PRESSED] = f.pressed; T[T_TWISTED] = f.twisted;
    T[T_V1] = f.v1; T[T_V2] = f.v2; T[T_B1] = f.b1; T[T_B2] = f.b2; T[T_BB] = f.bb; T[T_LIMIT] = f.limit;
    T[T_AL] = f.al; T[T_X] = f.x; T[T_Y] = f.y; T[T_LL] = f.ll; T[T_T1] = f.t1_out; T[T_T2] = f.t2_out;
    for (int c = 0; c < 3; c++) { T[T_U + c] = f.u[c]; T[T_DW + c] = f.dw[c]; }
    T[T_RN1] = f.rn1; T[T_RN2] = f.rn2; T[T_RMOST] = f.roll_most; T[T_R1OUT] = f.r1_out; T[T_R2OUT] = f.r2_out;
    T[T_SOUT] = f.s_out;
  }
  for (int k = 0; k < N; k++) {
    const int j = P * i + k;
    if (!known && !bias) W.now()[j] = in.sep[k];
    W.impulse()[j] = f.next[k];
    if ((f.first >> k) & 1) W.relvel()[j] = f.closing[k];
    if ((f.larger >> k) & 1) W.most()[j] = f.dd[k];
  }
  W.sliding[i] = f.slide ? 1 : 0;
  W.tangent()[2 * i] = f.t1_out; W.tangent()[2 * i + 1] = f.t2_out;
  if (N <= 2) { W.roll()[2 * i] = f.r1_out; W.roll()[2 * i + 1] = f.r2_out; }
  W.spin()[i] = f.s_out;
}

// KIND 0 warm start, 1 solve, 2 relax; the tape for this sweep's rows if given.
template <int KIND>
HD void row_of(const Bodies& B, const Work& W, const Rows& R, int boff, int r, int i, const Soft& soft, double h,
               bool known, const TapeRow& tape) {
  if (KIND == 0) { warm_row(B, W, R, boff, r, i); return; }
  switch (R.count[r]) {
    case 1: row_forward<1>(B, W, R, boff, r, i, soft, h, KIND == 1, known, tape); break;
    case 2: row_forward<2>(B, W, R, boff, r, i, soft, h, KIND == 1, known, tape); break;
    case 3: row_forward<3>(B, W, R, boff, r, i, soft, h, KIND == 1, known, tape); break;
    default: row_forward<4>(B, W, R, boff, r, i, soft, h, KIND == 1, known, tape); break;
  }
}

// src/world.js:1495-1530 (_accelerate) for a dynamic body.
HD void accelerate_body(const Bodies& B, int g, double h, int substeps, double gx, double gy, double gz) {
  const double ax = gx * h, ay = gy * h, az = gz * h;
  const double most = SPIN_TURN / (h * substeps);
  double *v = B.v + 3 * g, *w = B.w + 3 * g;
  const double sc = B.gravity_scale[g];
  v[0] += ax * sc; v[1] += ay * sc; v[2] += az * sc;
  if (B.damping[g] != 0) {
    const double k = 1 / (1 + h * B.damping[g]);
    v[0] *= k; v[1] *= k; v[2] *= k;
  }
  if (B.angular_damping[g] != 0) {
    const double k = 1 / (1 + h * B.angular_damping[g]);
    w[0] *= k; w[1] *= k; w[2] *= k;
  }
  if (B.capped[g]) {
    const double w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    note(B.branch ? B.branch + g : nullptr, 0x10u | (w2 > most * most ? 1u : 0u));
    if (w2 > most * most) {
      const double k = most / dsqrt(w2);
      w[0] *= k; w[1] *= k; w[2] *= k;
    }
  }
}

// src/world.js:1533-1563 (_move) for a dynamic body.
HD void move_body(const Bodies& B, int g, double h) {
  double *p = B.p + 3 * g, *v = B.v + 3 * g, *w = B.w + 3 * g;
  p[0] += v[0] * h; p[1] += v[1] * h; p[2] += v[2] * h;
  const bool turning = w[0] != 0 || w[1] != 0 || w[2] != 0;
  note(B.branch ? B.branch + g : nullptr, 0x20u | (turning ? 1u : 0u));
  if (turning && B.tumbles[g]) {
    double q[4], wo[3], c;
    unsigned took = 0;
    tumble_t<double>(B.q + 4 * g, w, B.left + 3 * g, B.ceiling[g], B.inertia + 3 * g, B.inv_inertia + 3 * g, h, q, wo,
                     c, &took);
    note(B.branch ? B.branch + g : nullptr, 0x40u | took);
    for (int k = 0; k < 4; k++) B.q[4 * g + k] = q[k];
    for (int k = 0; k < 3; k++) { w[k] = wo[k]; B.left[3 * g + k] = wo[k]; }
    B.ceiling[g] = c;
  } else if (turning) {
    double q[4];
    advance_t<double>(B.q + 4 * g, w, h * 0.5, q);
    for (int k = 0; k < 4; k++) B.q[4 * g + k] = q[k];
  }
}

// src/contact.js:2229-2251 (restitution) for one row, in narrow order.
HD void bounce_row(const Bodies& B, const Work& W, const Rows& R, int boff, int r, int i, double gravity) {
  const double e = W.e()[i];
  if (e == 0) return;
  const double threshold = dsqrt(2 * gravity * R.tolerance[r]);
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const int i3 = 3 * i;
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  for (int k = 0; k < R.count[r]; k++) {
    const int j = P * i + k;
    const bool skip = W.relvel()[j] > -threshold || W.most()[j] == 0;
    note(R.branch ? R.branch + r : nullptr, 0x400u | (skip ? 1u : 0u));
    if (skip) continue;
    const double* jac = W.jac() + J * j;
    Pair s = gather(B, ga, gb);
    const double vn = closing(s, nx, ny, nz, jac);
    const double old = W.impulse()[j];
    if (W.bounce()) {
      double* o = W.bounce() + BOUNCE * j;
      o[0] = old; o[1] = vn;
      for (int c = 0; c < 12; c++) o[2 + c] = s.v[c];
    }
    double next = old - W.nmass()[j] * (vn + e * W.relvel()[j]);
    note(R.branch ? R.branch + r : nullptr, 0x500u | (next < 0 ? 1u : 0u));
    if (next < 0) next = 0;
    const double l = next - old;
    W.impulse()[j] = next;
    push_normal(s, ima, imb, nx, ny, nz, jac, l);
    for (int c = 0; c < 3; c++) {
      B.v[3 * ga + c] = s.v[c]; B.w[3 * ga + c] = s.v[3 + c];
      B.v[3 * gb + c] = s.v[6 + c]; B.w[3 * gb + c] = s.v[9 + c];
    }
  }
}

// Whether row r's manifold was met at the moment of impact (its points carried back).
HD bool impacted(const Rows& R, int r) {
  for (int k = 0; k < R.count[r]; k++) {
    const double* sh = R.shift + 3 * (P * r + k);
    if (sh[0] != 0 || sh[1] != 0 || sh[2] != 0) return true;
  }
  return false;
}

// src/contact.js:1122-1153 (slipped), 1091-1115 (keep) for row r: its slip, then what
// its pair keeps for the next step.
HD void store_row(const Bodies& B, const Work& W, const Rows& R, int boff, int r, int i) {
  const int i3 = 3 * i, count = R.count[r];
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  double pressed = 0;
  for (int k = 0; k < count; k++) pressed += W.impulse()[P * i + k];
  const double t1x = W.t1()[i3], t1y = W.t1()[i3 + 1], t1z = W.t1()[i3 + 2];
  const double t2x = W.t2()[i3], t2y = W.t2()[i3 + 1], t2z = W.t2()[i3 + 2];
  const bool still = W.sliding[i] != 0 || pressed == 0 || impacted(R, r);
  note(R.branch ? R.branch + r : nullptr, 0x600u | (still ? 1u : 0u));
  if (still) {
    W.drift()[2 * i] = 0; W.drift()[2 * i + 1] = 0;
  } else {
    double T[6];
    double dx = 0, dy = 0, dz = 0;
    if (B.type[gb] == DYNAMIC) {
      turned_t<double>(B.p + 3 * gb, B.start + 3 * gb, B.q + 4 * gb, B.start_q + 4 * gb, T);
      const double rx = W.anchor_b()[i3], ry = W.anchor_b()[i3 + 1], rz = W.anchor_b()[i3 + 2];
      dx = T[0] + T[4] * rz - T[5] * ry;
      dy = T[1] + T[5] * rx - T[3] * rz;
      dz = T[2] + T[3] * ry - T[4] * rx;
    }
    if (B.type[ga] == DYNAMIC) {
      turned_t<double>(B.p + 3 * ga, B.start + 3 * ga, B.q + 4 * ga, B.start_q + 4 * ga, T);
      const double rx = W.anchor_a()[i3], ry = W.anchor_a()[i3 + 1], rz = W.anchor_a()[i3 + 2];
      dx -= T[0] + T[4] * rz - T[5] * ry;
      dy -= T[1] + T[5] * rx - T[3] * rz;
      dz -= T[2] + T[3] * ry - T[4] * rx;
    }
    W.drift()[2 * i] += dx * t1x + dy * t1y + dz * t1z;
    W.drift()[2 * i + 1] += dx * t2x + dy * t2y + dz * t2z;
  }
  const double nx = W.normal()[i3], ny = W.normal()[i3 + 1], nz = W.normal()[i3 + 2];
  const double r1 = W.roll()[2 * i], r2 = W.roll()[2 * i + 1], sp = W.spin()[i];
  double* L = R.store_twist + 3 * r;
  L[0] = t1x * r1 + t2x * r2 + nx * sp;
  L[1] = t1y * r1 + t2y * r2 + ny * sp;
  L[2] = t1z * r1 + t2z * r2 + nz * sp;
  const double l1 = W.tangent()[2 * i], l2 = W.tangent()[2 * i + 1];
  double* F = R.store_friction + 3 * r;
  F[0] = t1x * l1 + t2x * l2; F[1] = t1y * l1 + t2y * l2; F[2] = t1z * l1 + t2z * l2;
  const double d1 = W.drift()[2 * i], d2 = W.drift()[2 * i + 1];
  double* S = R.store_slip + 3 * r;
  S[0] = t1x * d1 + t2x * d2; S[1] = t1y * d1 + t2y * d2; S[2] = t1z * d1 + t2z * d2;
  for (int k = 0; k < count; k++) R.store_impulse[P * r + k] = W.impulse()[P * i + k];
}

// The tapes of one step: a row's inputs for each solve and relax sweep
// ([2 sub + relax][working row][ROW_TAPE]); each body's state before each substep and
// before each move ([2 sub + 1][body][STATE]).
struct Tape {
  double *rows, *bodies;
  int row_stride, body_stride;  // working rows and bodies a sweep's block holds
  HD TapeRow row(int sweep, int i) const {
    return TapeRow{rows ? rows + (size_t)sweep * ROW_TAPE * row_stride + i : nullptr, (size_t)row_stride};
  }
  HD double* body(int phase, int g) const { return bodies ? bodies + ((size_t)phase * body_stride + g) * STATE : nullptr; }
};

HD void save_state(const Bodies& B, int g, double* o) {
  for (int k = 0; k < 3; k++) o[k] = B.p[3 * g + k];
  for (int k = 0; k < 4; k++) o[3 + k] = B.q[4 * g + k];
  for (int k = 0; k < 3; k++) o[7 + k] = B.v[3 * g + k];
  for (int k = 0; k < 3; k++) o[10 + k] = B.w[3 * g + k];
  for (int k = 0; k < 3; k++) o[13 + k] = B.left[3 * g + k];
  o[16] = B.ceiling[g];
}
HD void load_state(const Bodies& B, int g, const double* o) {
  for (int k = 0; k < 3; k++) B.p[3 * g + k] = o[k];
  for (int k = 0; k < 4; k++) B.q[4 * g + k] = o[3 + k];
  for (int k = 0; k < 3; k++) B.v[3 * g + k] = o[7 + k];
  for (int k = 0; k < 3; k++) B.w[3 * g + k] = o[10 + k];
  for (int k = 0; k < 3; k++) B.left[3 * g + k] = o[13 + k];
  B.ceiling[g] = o[16];
}

// One sweep over a world's colours, a barrier after each, then its overflow in order.
template <int KIND, class Sync>
HD void sweep(const Ctx& c, Sync sync, const Bodies& B, const Work& W, const Rows& R, const StepView& v,
              const Soft& soft, double h, bool known, const Tape& tape, int tsweep) {
  for (int col = 0; col < COLOURS; col++) {
    const int from = v.starts[col], to = v.starts[col + 1];
    if (from == to) continue;
    for (int lm = from + c.tid; lm < to; lm += c.threads) {
      const int i = v.w0 + lm;
      row_of<KIND>(B, W, R, v.boff, v.r0 + lm, i, soft, h, known, tape.row(KIND == 0 ? 0 : tsweep, i));
    }
    sync();
  }
  const int from = v.starts[OVERFLOW], to = v.starts[OVERFLOW + 1];
  if (from < to) {
    if (c.tid == 0)
      for (int lm = from; lm < to; lm++) {
        const int i = v.w0 + lm;
        row_of<KIND>(B, W, R, v.boff, v.r0 + lm, i, soft, h, known, tape.row(KIND == 0 ? 0 : tsweep, i));
      }
    sync();
  }
}

// A held step of one world (src/world.js:961-1031 with each manifold placed from its
// bodies first); `placed`: the caller filled the rows' places and impulses itself.
template <class Sync>
HD void step_forward(const Ctx& c, Sync sync, const Bodies& B, const Sizes& Z, const Rows& R, const Work& W,
                     const StepView& v, const Tape& tape, bool placed) {
  const int nrows = v.r1 - v.r0;
  if (!placed) {
    for (int lm = c.tid; lm < nrows; lm += c.threads) {
      place_row(B, R, W, v.boff, v.r0 + lm, v.w0 + lm);
      carry_row(R, W, v.r0 + lm, v.w0 + lm);
    }
  }
  for (int s = c.tid; s < v.nbodies; s += c.threads) {
    const int g = v.boff + s;
    inertia_body(B, g);
    for (int k = 0; k < 3; k++) B.start[3 * g + k] = B.p[3 * g + k];
    for (int k = 0; k < 4; k++) B.start_q[4 * g + k] = B.q[4 * g + k];
  }
  sync();
  const int substeps = v.substeps;
  const double h = v.dt / substeps;
  const bool touching = nrows > 0;
  Soft soft{0, 1, 0};
  if (touching) {
    for (int lm = c.tid; lm < nrows; lm += c.threads) prepare_row(B, Z, W, R, v.boff, v.r0 + lm, v.w0 + lm);
    soft = make_soft(STIFFNESS / h, DAMPING_RATIO, h);
    sync();
  }
  for (int sub = 0; sub < substeps; sub++) {
    for (int s = c.tid; s < v.nbodies; s += c.threads) {
      const int g = v.boff + s;
      if (B.type[g] != DYNAMIC) continue;
      if (tape.bodies) save_state(B, g, tape.body(2 * sub, g));
      accelerate_body(B, g, h, substeps, v.gx, v.gy, v.gz);
    }
    sync();
    if (touching) {
      sweep<0>(c, sync, B, W, R, v, soft, h, false, tape, 0);
      sweep<1>(c, sync, B, W, R, v, soft, h, sub > 0, tape, 2 * sub);
    }
    for (int s = c.tid; s < v.nbodies; s += c.threads) {
      const int g = v.boff + s;
      if (B.type[g] != DYNAMIC) continue;
      if (tape.bodies) save_state(B, g, tape.body(2 * sub + 1, g));
      move_body(B, g, h);
    }
    sync();
    if (touching) sweep<2>(c, sync, B, W, R, v, soft, h, false, tape, 2 * sub + 1);
  }
  if (touching) {
    if (c.tid == 0) {
      const double gravity = dsqrt(v.gx * v.gx + v.gy * v.gy + v.gz * v.gz);
      for (int f = 0; f < nrows; f++) {
        const int lm = v.narrow[f];
        bounce_row(B, W, R, v.boff, v.r0 + lm, v.w0 + lm, gravity);
      }
    }
    sync();
    for (int lm = c.tid; lm < nrows; lm += c.threads) store_row(B, W, R, v.boff, v.r0 + lm, v.w0 + lm);
    sync();
  }
}


// ---------------------------------------------------------------------------------
// The reverse.

// Adjoints by body (global slot). Only bodies the step moves (dynamic) carry any.
struct BodiesBar {
  double *p, *q, *v, *w, *left, *ceiling;  // of the state at the reverse's current point
  double *inv_i, *start, *start_q;         // of the step's own
  double *inv_mass, *inertia, *inv_inertia;  // of the mass-derived columns, summed over every step
};

// Adjoints by working row: of the row's state, of what prepare derived, of its places, and
// what its sweeps added to its bodies' inverse mass and inertia adjoints.
struct WorkBar {
  double* base;  // the columns, one after another, each n rows wide
  size_t n;
  HD double* impulse() const { return base + 0 * n; }  // 4 a row
  HD double* tangent() const { return base + 4 * n; }  // 2 a row
  HD double* roll() const { return base + 6 * n; }  // 2 a row
  HD double* spin() const { return base + 8 * n; }  // 1 a row
  HD double* drift() const { return base + 9 * n; }  // 2 a row
  HD double* relvel() const { return base + 11 * n; }  // 4 a row
  HD double* jac() const { return base + 15 * n; }  // 48 a row
  HD double* fjac() const { return base + 63 * n; }  // 24 a row
  HD double* tm() const { return base + 87 * n; }  // 3 a row
  HD double* nmass() const { return base + 90 * n; }  // 4 a row
  HD double* block() const { return base + 94 * n; }  // 16 a row
  HD double* rm() const { return base + 110 * n; }  // 2 a row
  HD double* sm() const { return base + 112 * n; }  // 1 a row
  HD double* lever() const { return base + 113 * n; }  // 4 a row
  HD double* la() const { return base + 117 * n; }  // 12 a row
  HD double* lb() const { return base + 129 * n; }  // 12 a row
  HD double* adjusted() const { return base + 141 * n; }  // 4 a row
  HD double* normal() const { return base + 145 * n; }  // 3 a row
  HD double* t1() const { return base + 148 * n; }  // 3 a row
  HD double* t2() const { return base + 151 * n; }  // 3 a row
  HD double* point() const { return base + 154 * n; }  // 12 a row
  HD double* separation() const { return base + 166 * n; }  // 4 a row
  HD double* mu() const { return base + 170 * n; }  // 1 a row
  HD double* e() const { return base + 171 * n; }  // 1 a row
  HD double* anchor_a() const { return base + 172 * n; }  // 3 a row
  HD double* anchor_b() const { return base + 175 * n; }  // 3 a row
  HD double* ima() const { return base + 178 * n; }  // 1 a row
  HD double* imb() const { return base + 179 * n; }  // 1 a row
  HD double* Ia() const { return base + 180 * n; }  // 6 a row
  HD double* Ib() const { return base + 186 * n; }  // 6 a row
  static constexpr int WIDTH = 192;
};

// Adjoints by global row: the stored impulses, and mu and e (summed onto the bodies'
// friction and restitution at the end).
struct RowsBar {
  double *store_impulse, *store_friction, *store_twist, *store_slip, *mu, *e;
};

HD bool moves(const Bodies& B, int g) { return B.type[g] == DYNAMIC; }

// One working row's adjoints (its WorkBar fields but the anchors' and adjusted separations')
// while a sweep's reverse works on the row, in the thread's slot (Ctx::slot: LDS on the
// device): read once before, written once after. Between the two the row's reverse adds to
// them in the slot and stores nothing to global memory until its separations, so its reads of
// the tape, the forward's columns and the bodies need not wait on stores before them. A warm
// start's reverse (SOLVE false) touches only the first group. Every sum is taken in the same
// order as on the columns themselves.
template <int N, bool SOLVE>
struct RowAcc {
  double impulse[N], tangent[2], roll[2], spin, jac[N * J], fjac[FJ], normal[3], t1[3], t2[3], ima, imb, Ia[6], Ib[6];
  double drift[2], relvel[N], tm[3], nmass[N], block[P * P], rm[2], sm, lever[N], mu;

  HD void load(const WorkBar& WB, int i) {
    for (int k = 0; k < N; k++) impulse[k] = WB.impulse()[P * i + k];
    for (int c = 0; c < 2; c++) { tangent[c] = WB.tangent()[2 * i + c]; roll[c] = WB.roll()[2 * i + c]; }
    spin = WB.spin()[i];
    for (int c = 0; c < N * J; c++) jac[c] = WB.jac()[(size_t)J * P * i + c];
    for (int c = 0; c < FJ; c++) fjac[c] = WB.fjac()[(size_t)FJ * i + c];
    for (int c = 0; c < 3; c++) { normal[c] = WB.normal()[3 * i + c]; t1[c] = WB.t1()[3 * i + c]; t2[c] = WB.t2()[3 * i + c]; }
    ima = WB.ima()[i];
    imb = WB.imb()[i];
    for (int c = 0; c < 6; c++) { Ia[c] = WB.Ia()[6 * i + c]; Ib[c] = WB.Ib()[6 * i + c]; }
    if (!SOLVE) return;
    for (int c = 0; c < 2; c++) { drift[c] = WB.drift()[2 * i + c]; rm[c] = WB.rm()[2 * i + c]; }
    for (int k = 0; k < N; k++) {
      relvel[k] = WB.relvel()[P * i + k]; nmass[k] = WB.nmass()[P * i + k];
      lever[k] = WB.lever()[P * i + k];
    }
    for (int c = 0; c < 3; c++) tm[c] = WB.tm()[3 * i + c];
    for (int c = 0; c < P * P; c++) block[c] = WB.block()[P * P * i + c];
    sm = WB.sm()[i];
    mu = WB.mu()[i];
  }
  HD void store(const WorkBar& WB, int i) const {
    for (int k = 0; k < N; k++) WB.impulse()[P * i + k] = impulse[k];
    for (int c = 0; c < 2; c++) { WB.tangent()[2 * i + c] = tangent[c]; WB.roll()[2 * i + c] = roll[c]; }
    WB.spin()[i] = spin;
    for (int c = 0; c < N * J; c++) WB.jac()[(size_t)J * P * i + c] = jac[c];
    for (int c = 0; c < FJ; c++) WB.fjac()[(size_t)FJ * i + c] = fjac[c];
    for (int c = 0; c < 3; c++) { WB.normal()[3 * i + c] = normal[c]; WB.t1()[3 * i + c] = t1[c]; WB.t2()[3 * i + c] = t2[c]; }
    WB.ima()[i] = ima;
    WB.imb()[i] = imb;
    for (int c = 0; c < 6; c++) { WB.Ia()[6 * i + c] = Ia[c]; WB.Ib()[6 * i + c] = Ib[c]; }
    if (!SOLVE) return;
    for (int c = 0; c < 2; c++) { WB.drift()[2 * i + c] = drift[c]; WB.rm()[2 * i + c] = rm[c]; }
    for (int k = 0; k < N; k++) {
      WB.relvel()[P * i + k] = relvel[k]; WB.nmass()[P * i + k] = nmass[k];
      WB.lever()[P * i + k] = lever[k];
    }
    for (int c = 0; c < 3; c++) WB.tm()[3 * i + c] = tm[c];
    for (int c = 0; c < P * P; c++) WB.block()[P * P * i + c] = block[c];
    WB.sm()[i] = sm;
    WB.mu()[i] = mu;
  }
};
// A slot's doubles: room for the largest RowAcc; odd, so the slots of a wave's lanes start in
// different LDS banks.
constexpr int ROW_SLOT = 141;
static_assert(sizeof(RowAcc<4, true>) <= ROW_SLOT * sizeof(double), "a row's adjoints fit a slot");

// y = I x for a symmetric I (6: xx xy xz yy yz zz): adds the adjoints of I and x.
HD void sym_vjp(const double* I, const double* x, const double* yb, double* Ib, double* xb) {
  if (Ib) {
    Ib[0] += yb[0] * x[0];
    Ib[1] += yb[0] * x[1] + yb[1] * x[0];
    Ib[2] += yb[0] * x[2] + yb[2] * x[0];
    Ib[3] += yb[1] * x[1];
    Ib[4] += yb[1] * x[2] + yb[2] * x[1];
    Ib[5] += yb[2] * x[2];
  }
  if (xb) {
    xb[0] += I[0] * yb[0] + I[1] * yb[1] + I[2] * yb[2];
    xb[1] += I[1] * yb[0] + I[3] * yb[1] + I[4] * yb[2];
    xb[2] += I[2] * yb[0] + I[4] * yb[1] + I[5] * yb[2];
  }
}

// The reverse of push_normal: s += (-ima n l, -jac[6..8] l, imb n l, jac[9..11] l).
// Adds the adjoints of l, n, ima, imb and jac[6..11] from the velocities' adjoints sb.
HD double push_normal_vjp(const Pair& sb, double ima, double imb, const double* n, const double* jac, double l,
                          double* nb, double& imab, double& imbb, double* jacb) {
  double lb = 0;
  for (int c = 0; c < 3; c++) {
    const double d = imb * sb.v[6 + c] - ima * sb.v[c];
    lb += n[c] * d + jac[9 + c] * sb.v[9 + c] - jac[6 + c] * sb.v[3 + c];
    nb[c] += l * d;
    imab -= n[c] * l * sb.v[c];
    imbb += n[c] * l * sb.v[6 + c];
    jacb[6 + c] -= l * sb.v[3 + c];
    jacb[9 + c] += l * sb.v[9 + c];
  }
  return lb;
}

// The reverse of push_turn(L): wa -= Ia L, wb += Ib L. Returns L's adjoint in Lb.
HD void push_turn_vjp(const Pair& sb, const double* Ia, const double* Ib, const double* L, double* Lb, double* Iab,
                      double* Ibb) {
  const double wa[3] = {-sb.v[3], -sb.v[4], -sb.v[5]};
  const double wb[3] = {sb.v[9], sb.v[10], sb.v[11]};
  sym_vjp(Ia, L, wa, Iab, Lb);
  sym_vjp(Ib, L, wb, Ibb, Lb);
}

// The reverse of closing(s, n, jac) with adjoint cb: into sb, n and jac[0..5].
HD void closing_vjp(const Pair& s, const double* n, const double* jac, double cb, Pair& sb, double* nb, double* jacb) {
  for (int c = 0; c < 3; c++) {
    sb.v[6 + c] += cb * n[c];
    sb.v[c] -= cb * n[c];
    sb.v[9 + c] += cb * jac[3 + c];
    sb.v[3 + c] -= cb * jac[c];
    nb[c] += cb * (s.v[6 + c] - s.v[c]);
    jacb[3 + c] += cb * s.v[9 + c];
    jacb[c] -= cb * s.v[3 + c];
  }
}

// x = B^-1 n with B = I + al M, M = [[a, b], [b, c]] (2x2, symmetric).
HD void solve2(double al, double a, double b, double c, double n1, double n2, double& x, double& y) {
  const double ab = al * b, det = (1 + al * a) * (1 + al * c) - ab * ab;
  x = ((1 + al * c) * n1 - ab * n2) / det;
  y = ((1 + al * a) * n2 - ab * n1) / det;
}

// The reverse of a solve (bias) or relax row of N points, from what its forward left on
// the tape (T, the fields below) rather than by working the row again. sb holds its bodies'
// velocity adjoints after the row and leaves with them before it; the row's state adjoints
// in WB go from after to before; what the row read of prepare, the normal and the bodies'
// inverse masses and inertias gets its adjoints added; sepb gets the separations'.
template <int N>
HD void solve_vjp(const Work& W, RowAcc<N, true>& A, const Rows& R, int r, int i, double ima, double imb,
                  const double* Ia, const double* Ib, const Soft& soft, double h, bool bias, const TapeRow& T,
                  Pair& sb, double& imab, double& imbb, double (&sepb)[N]) {
  double* Iab = A.Ia;
  double* Ibb = A.Ib;
  const int i3 = 3 * i;
  // The row's own columns read at once, before the reverse's first use of any.
  double n[3], t1[3], t2[3], tmv[3], lever_in[N], nmass_in[N];
  for (int c = 0; c < 3; c++) { n[c] = W.normal()[i3 + c]; t1[c] = W.t1()[i3 + c]; t2[c] = W.t2()[i3 + c]; tmv[c] = W.tm()[i3 + c]; }
  for (int k = 0; k < N; k++) { lever_in[k] = W.lever()[P * i + k]; nmass_in[k] = W.nmass()[P * i + k]; }
  const double sm_in = W.sm()[i], rm1_in = W.rm()[2 * i], rm2_in = W.rm()[2 * i + 1], tol_in = R.tolerance[r];
  double* nb = A.normal;
  double* t1b = A.t1;
  double* t2b = A.t2;
  const double mu = W.mu()[i], patch = R.patch[r];
  const uint32_t flags = (uint32_t)T[T_FLAGS];
  const double pressed = T[T_PRESSED], twisted = T[T_TWISTED];
  double pressed_b = 0, twisted_b = 0;
  // spin
  {
    const double s0 = T[T_S0], ds = T[T_SOUT] - s0;
    const double L[3] = {n[0] * ds, n[1] * ds, n[2] * ds};
    double Lb[3] = {0, 0, 0};
    push_turn_vjp(sb, Ia, Ib, L, Lb, Iab, Ibb);
    const double dsb = n[0] * Lb[0] + n[1] * Lb[1] + n[2] * Lb[2];
    for (int c = 0; c < 3; c++) nb[c] += ds * Lb[c];
    const double s1b = A.spin + dsb;
    double s0b = -dsb;
    double limitb = 0;
    const int sclamp = (flags >> 3) & 3;
    if (sclamp == 1) limitb += s1b;
    else if (sclamp == 2) limitb -= s1b;
    else {
      s0b += s1b;
      const double dw[3] = {T[T_DW], T[T_DW + 1], T[T_DW + 2]};
      const double proj = dw[0] * n[0] + dw[1] * n[1] + dw[2] * n[2];
      A.sm -= s1b * proj;
      const double sm = sm_in;
      for (int c = 0; c < 3; c++) {
        const double g = -s1b * sm * n[c];
        sb.v[9 + c] += g; sb.v[3 + c] -= g;
        nb[c] -= s1b * sm * dw[c];
      }
    }
    if (N == 1) {
      A.mu += SPIN_DISC * patch * pressed * limitb;
      pressed_b += SPIN_DISC * mu * patch * limitb;
    } else {
      A.mu += twisted * limitb;
      twisted_b += mu * limitb;
    }
    A.spin = s0b;
  }
  // roll
  if (N <= 2) {
    const double r1 = T[T_R1], r2 = T[T_R2];
    const double d1 = T[T_R1OUT] - r1, d2 = T[T_R2OUT] - r2;
    const double L[3] = {t1[0] * d1 + t2[0] * d2, t1[1] * d1 + t2[1] * d2, t1[2] * d1 + t2[2] * d2};
    double Lb[3] = {0, 0, 0};
    push_turn_vjp(sb, Ia, Ib, L, Lb, Iab, Ibb);
    const double d1b = t1[0] * Lb[0] + t1[1] * Lb[1] + t1[2] * Lb[2];
    const double d2b = t2[0] * Lb[0] + t2[1] * Lb[1] + t2[2] * Lb[2];
    for (int c = 0; c < 3; c++) { t1b[c] += d1 * Lb[c]; t2b[c] += d2 * Lb[c]; }
    const double o1b = A.roll[0] + d1b, o2b = A.roll[1] + d2b;
    double r1b = -d1b, r2b = -d2b;
    double n1b = o1b, n2b = o2b;
    if ((flags >> 2) & 1) {
      // out = most * n / |n|
      const double rn1 = T[T_RN1], rn2 = T[T_RN2];
      const double l = dsqrt(rn1 * rn1 + rn2 * rn2);
      const double u1 = rn1 / l, u2 = rn2 / l;
      const double ou = o1b * u1 + o2b * u2;
      pressed_b += patch * ou;
      const double k = T[T_RMOST] / l;
      n1b = k * (o1b - u1 * ou);
      n2b = k * (o2b - u2 * ou);
    }
    r1b += n1b; r2b += n2b;
    const double rm1 = rm1_in, rm2 = rm2_in;
    const double u[3] = {T[T_U], T[T_U + 1], T[T_U + 2]};
    const double p1 = u[0] * t1[0] + u[1] * t1[1] + u[2] * t1[2];
    const double p2 = u[0] * t2[0] + u[1] * t2[1] + u[2] * t2[2];
    A.rm[0] -= n1b * p1;
    A.rm[1] -= n2b * p2;
    for (int c = 0; c < 3; c++) {
      const double ub = -n1b * rm1 * t1[c] - n2b * rm2 * t2[c];
      sb.v[9 + c] += ub; sb.v[3 + c] -= ub;
      t1b[c] -= n1b * rm1 * u[c];
      t2b[c] -= n2b * rm2 * u[c];
    }
    A.roll[0] = r1b; A.roll[1] = r2b;
  }
  // friction
  {
    const double* fj = W.fjac() + FJ * i;
    const double* g = fj + 12;
    double* fjb = A.fjac;
    double* gb = fjb + 12;
    const double o1 = T[T_O1], o2 = T[T_O2];
    const double d1 = T[T_T1] - o1, d2 = T[T_T2] - o2;
    double d1b = 0, d2b = 0;
    for (int c = 0; c < 3; c++) {
      const double dv = imb * sb.v[6 + c] - ima * sb.v[c];
      d1b += t1[c] * dv + fj[9 + c] * sb.v[9 + c] - fj[6 + c] * sb.v[3 + c];
      d2b += t2[c] * dv + g[9 + c] * sb.v[9 + c] - g[6 + c] * sb.v[3 + c];
      t1b[c] += d1 * dv;
      t2b[c] += d2 * dv;
      const double p = t1[c] * d1 + t2[c] * d2;
      imab -= p * sb.v[c];
      imbb += p * sb.v[6 + c];
      fjb[6 + c] -= d1 * sb.v[3 + c]; gb[6 + c] -= d2 * sb.v[3 + c];
      fjb[9 + c] += d1 * sb.v[9 + c]; gb[9 + c] += d2 * sb.v[9 + c];
    }
    const double o1b = A.tangent[0] + d1b, o2b = A.tangent[1] + d2b;
    double in1b = -d1b, in2b = -d2b;  // of the tangent impulses before
    double n1b = o1b, n2b = o2b;      // of n before the cone
    const double tm0 = tmv[0], tm1 = tmv[1], tm2 = tmv[2];
    double* tmb = A.tm;
    if ((flags >> 1) & 1) {
      // out = limit x / |x|, x = (I + al M)^-1 n, al where |x| = limit.
      const double limit = T[T_LIMIT], al = T[T_AL], fx = T[T_X], fy = T[T_Y];
      double limitb = 0;
      const double lx = dsqrt(T[T_LL]);
      const double u1 = fx / lx, u2 = fy / lx;
      const double ou = o1b * u1 + o2b * u2;
      limitb += ou;
      const double k = limit / lx;
      double xb = k * (o1b - u1 * ou), yb = k * (o2b - u2 * ou);
      n1b = 0; n2b = 0;
      if (limit > 0 && al > 0) {
        double bx, by, bmx, bmy;
        solve2(al, tm0, tm1, tm2, xb, yb, bx, by);
        const double mx = tm0 * fx + tm1 * fy, my = tm1 * fx + tm2 * fy;
        solve2(al, tm0, tm1, tm2, mx, my, bmx, bmy);
        const double alb = -(bx * mx + by * my);
        const double gal = -2 * (fx * bmx + fy * bmy);
        const double cc = gal != 0 ? -alb / gal : 0;
        const double zx = xb + 2 * cc * fx, zy = yb + 2 * cc * fy;
        double yx, yy;
        solve2(al, tm0, tm1, tm2, zx, zy, yx, yy);
        n1b += yx; n2b += yy;
        // M's adjoint: -al y x^T
        tmb[0] -= al * yx * fx;
        tmb[1] -= al * (yx * fy + yy * fx);
        tmb[2] -= al * yy * fy;
        limitb += cc * (-2 * limit);
      } else {
        // limit 0, or no Newton step taken: x = n
        n1b += xb; n2b += yb;
      }
      A.mu += pressed * limitb;
      pressed_b += mu * limitb;
    }
    // n = o - M v
    in1b += n1b; in2b += n2b;
    const double v1 = T[T_V1], v2 = T[T_V2];
    const double v1b = -(tm0 * n1b + tm1 * n2b), v2b = -(tm1 * n1b + tm2 * n2b);
    tmb[0] -= n1b * v1;
    tmb[1] -= n1b * v2 + n2b * v1;
    tmb[2] -= n2b * v2;
    // the slip's pull (bias) depends only on the drift
    const int fbias = (flags >> 5) & 3;
    if (fbias == 1) {
      A.drift[0] += soft.bias_rate * v1b;
      A.drift[1] += soft.bias_rate * v2b;
    } else if (fbias == 2) {
      const double tol = tol_in, max_bias = tol / h;
      const double l = dsqrt(T[T_BB]);
      const double u1 = T[T_B1] / l, u2 = T[T_B2] / l, k = max_bias / l;
      const double ou = v1b * u1 + v2b * u2;
      A.drift[0] += soft.bias_rate * k * (v1b - u1 * ou);
      A.drift[1] += soft.bias_rate * k * (v2b - u2 * ou);
    }
    // v1, v2 from the velocities after the normal pushes
    for (int c = 0; c < 3; c++) {
      const double ma = T[T_MID + c], mwa = T[T_MID + 3 + c], mb = T[T_MID + 6 + c], mwb = T[T_MID + 9 + c];
      const double dv = mb - ma;
      t1b[c] += v1b * dv; t2b[c] += v2b * dv;
      const double gv = v1b * t1[c] + v2b * t2[c];
      sb.v[6 + c] += gv; sb.v[c] -= gv;
      sb.v[9 + c] += v1b * fj[3 + c] + v2b * g[3 + c];
      sb.v[3 + c] -= v1b * fj[c] + v2b * g[c];
      fjb[3 + c] += v1b * mwb; gb[3 + c] += v2b * mwb;
      fjb[c] -= v1b * mwa; gb[c] -= v2b * mwa;
    }
    A.tangent[0] = in1b; A.tangent[1] = in2b;
  }
  // normal
  double nextb[N], oldb[N];
  for (int k = 0; k < N; k++) {
    nextb[k] = A.impulse[k] + pressed_b + lever_in[k] * twisted_b;
    A.lever[k] += T[T_NEXT + k] * twisted_b;
    oldb[k] = 0;
  }
  for (int k = N - 1; k >= 0; k--) {
    const int j = P * i + k;
    const double ddb = push_normal_vjp(sb, ima, imb, n, W.jac() + J * j, T[T_DD + k], nb, imab, imbb, A.jac + J * k);
    nextb[k] += ddb;
    oldb[k] -= ddb;
  }
  const uint32_t first = (flags >> 12) & 15;
  double closingb[N], targetb[N];
  for (int k = 0; k < N; k++) {
    closingb[k] = 0; targetb[k] = 0;
    if ((first >> k) & 1) {
      closingb[k] += A.relvel[k];
      A.relvel[k] = 0;
    }
  }
  // what each point's target case made of its mass and impulse scales and give
  const double springy = soft.impulse_scale / soft.mass_scale + COMPLIANCE;
  auto tcase = [&](int k) { return (int)((flags >> (16 + 2 * k)) & 3); };
  if (flags & 1) {
    for (int k = 0; k < N; k++) {
      if (!(T[T_RAW + k] > 0)) continue;  // clamped at 0 (or exactly 0)
      const double g = nextb[k];
      const double nm = nmass_in[k];
      const bool soft_k = tcase(k) == 1 || tcase(k) == 2;
      const double ms = soft_k ? soft.mass_scale : 1, is = soft_k ? soft.impulse_scale : 0;
      oldb[k] += g * (1 - is);
      A.nmass[k] -= g * ms * (T[T_CLOSING + k] + T[T_TARGET + k]);
      closingb[k] -= g * nm * ms;
      targetb[k] -= g * nm * ms;
    }
  } else {
    // solved on the set S: A_SS x = wanted_S, A = K with diagonal scaled by 1 + give
    const double* Kg = W.block() + P * P * i;
    double* Kb = A.block;
    const int set = (flags >> 8) & 15;
    double xb[N], give[N], solved[N];
    int idx[N], ns = 0;
    for (int k = 0; k < N; k++) {
      solved[k] = T[T_SOLVED + k];
      const bool soft_k = tcase(k) == 1 || tcase(k) == 2;
      give[k] = soft_k ? springy : COMPLIANCE;
      xb[k] = (solved[k] > 0 && ((set >> k) & 1)) ? nextb[k] : 0;
      if ((set >> k) & 1) idx[ns++] = k;
    }
    // y = A_SS^-1 xb_S (A symmetric), by Gaussian elimination on the set
    double M[N][N + 1];
    for (int a = 0; a < ns; a++) {
      for (int b = 0; b < ns; b++) {
        const int ka = idx[a], kb = idx[b];
        M[a][b] = Kg[P * ka + kb] * (ka == kb ? 1 + give[ka] : 1);
      }
      M[a][ns] = xb[idx[a]];
    }
    for (int c = 0; c < ns; c++) {
      for (int r2 = c + 1; r2 < ns; r2++) {
        const double q = M[r2][c] / M[c][c];
        for (int c2 = c; c2 <= ns; c2++) M[r2][c2] -= q * M[c][c2];
      }
    }
    double y[N];
    for (int a = ns - 1; a >= 0; a--) {
      double v = M[a][ns];
      for (int b = a + 1; b < ns; b++) v -= M[a][b] * y[b];
      y[a] = v / M[a][a];
    }
    // wanted_S's adjoint is y; A's is -y x^T
    double wantedb[N];
    for (int k = 0; k < N; k++) wantedb[k] = 0;
    for (int a = 0; a < ns; a++) {
      wantedb[idx[a]] = y[a];
      for (int b = 0; b < ns; b++) {
        const int ka = idx[a], kb = idx[b];
        const double g = -y[a] * solved[kb];
        Kb[P * ka + kb] += ka == kb ? g * (1 + give[ka]) : g;
      }
    }
    // wanted_k = -closing_k - target_k + sum_l K_kl impulse_l
    for (int k = 0; k < N; k++) {
      const double g = wantedb[k];
      if (g == 0) continue;
      closingb[k] -= g;
      targetb[k] -= g;
      for (int l = 0; l < N; l++) {
        Kb[P * k + l] += g * T[T_IMP + l];
        oldb[l] += g * Kg[P * k + l];
      }
    }
  }
  Pair s0;
  for (int c = 0; c < 12; c++) s0.v[c] = T[T_S + c];
  for (int k = N - 1; k >= 0; k--) {
    const int j = P * i + k;
    closing_vjp(s0, n, W.jac() + J * j, closingb[k], sb, nb, A.jac + J * k);
    A.impulse[k] = oldb[k];
    // target(s)
    const double inv = 1 / h;
    double g = 0;
    if (tcase(k) == 0) g = targetb[k] * inv;
    else if (tcase(k) == 1) g = targetb[k] * soft.bias_rate;
    sepb[k] = g;
  }
}


// The reverse of warm_row: the velocities only gain pushes, so their adjoints pass
// through; the impulses, normal, tangents, lever terms and inertias gain theirs. The
// impulses are the row's as the solve that follows read them (its tape).
template <int N>
HD void warm_vjp(const Work& W, RowAcc<N, false>& A, const Rows& R, int r, int i, double ima, double imb,
                 const double* Ia, const double* Ib, const TapeRow& T, const Pair& sb, double& imab, double& imbb) {
  double* Iab = A.Ia;
  double* Ibb = A.Ib;
  const int count = R.count[r], i3 = 3 * i;
  // the row's columns and tape read at once, before the first use of any
  double n[3], t1[3], t2[3], imp[N], fj[FJ];
  for (int c = 0; c < 3; c++) { n[c] = W.normal()[i3 + c]; t1[c] = W.t1()[i3 + c]; t2[c] = W.t2()[i3 + c]; }
  for (int k = 0; k < N; k++) imp[k] = T[T_IMP + k];
  for (int c = 0; c < FJ; c++) fj[c] = W.fjac()[FJ * i + c];
  const double l1 = T[T_O1], l2 = T[T_O2], r1 = T[T_R1], r2 = T[T_R2], sp = T[T_S0];
  double* nb = A.normal;
  double* t1b = A.t1;
  double* t2b = A.t2;
  for (int k = 0; k < N; k++) {
    if (k >= count) break;
    const int j = P * i + k;
    A.impulse[k] += push_normal_vjp(sb, ima, imb, n, W.jac() + J * j, imp[k], nb, imab, imbb, A.jac + J * k);
  }
  {
    double* fjb = A.fjac;
    double l1b = 0, l2b = 0;
    for (int c = 0; c < 3; c++) {
      const double dv = imb * sb.v[6 + c] - ima * sb.v[c];
      l1b += t1[c] * dv + fj[9 + c] * sb.v[9 + c] - fj[6 + c] * sb.v[3 + c];
      l2b += t2[c] * dv + fj[21 + c] * sb.v[9 + c] - fj[18 + c] * sb.v[3 + c];
      t1b[c] += l1 * dv; t2b[c] += l2 * dv;
      const double p = t1[c] * l1 + t2[c] * l2;
      imab -= p * sb.v[c];
      imbb += p * sb.v[6 + c];
      fjb[6 + c] -= l1 * sb.v[3 + c]; fjb[18 + c] -= l2 * sb.v[3 + c];
      fjb[9 + c] += l1 * sb.v[9 + c]; fjb[21 + c] += l2 * sb.v[9 + c];
    }
    A.tangent[0] += l1b; A.tangent[1] += l2b;
  }
  {
    const double L[3] = {t1[0] * r1 + t2[0] * r2 + n[0] * sp, t1[1] * r1 + t2[1] * r2 + n[1] * sp,
                         t1[2] * r1 + t2[2] * r2 + n[2] * sp};
    double Lb[3] = {0, 0, 0};
    push_turn_vjp(sb, Ia, Ib, L, Lb, Iab, Ibb);
    if (count <= 2) {
      A.roll[0] += t1[0] * Lb[0] + t1[1] * Lb[1] + t1[2] * Lb[2];
      A.roll[1] += t2[0] * Lb[0] + t2[1] * Lb[1] + t2[2] * Lb[2];
    }
    A.spin += n[0] * Lb[0] + n[1] * Lb[1] + n[2] * Lb[2];
    for (int c = 0; c < 3; c++) { t1b[c] += r1 * Lb[c]; t2b[c] += r2 * Lb[c]; nb[c] += sp * Lb[c]; }
  }
}

// q's and v's adjoints of rotate(q, v, inverse) from the output's: with u = +-q's vector
// part, t = 2 u x v and out = v + w t + u x t, t's adjoint is w ob + ob x u, u's gains
// t x ob and 2 v x tb, v's ob and 2 tb x u, w's ob . t.
HD void rotate_vjp(const double* q, const double* v, bool inverse, const double* ob, double* qb, double* vb) {
  const double s = inverse ? -1 : 1;
  const double u[3] = {q[0] * s, q[1] * s, q[2] * s}, w = q[3];
  const double t[3] = {2 * (u[1] * v[2] - u[2] * v[1]), 2 * (u[2] * v[0] - u[0] * v[2]), 2 * (u[0] * v[1] - u[1] * v[0])};
  const double tb[3] = {w * ob[0] + (ob[1] * u[2] - ob[2] * u[1]), w * ob[1] + (ob[2] * u[0] - ob[0] * u[2]),
                        w * ob[2] + (ob[0] * u[1] - ob[1] * u[0])};
  if (qb) {
    const double ub[3] = {(t[1] * ob[2] - t[2] * ob[1]) + 2 * (v[1] * tb[2] - v[2] * tb[1]),
                          (t[2] * ob[0] - t[0] * ob[2]) + 2 * (v[2] * tb[0] - v[0] * tb[2]),
                          (t[0] * ob[1] - t[1] * ob[0]) + 2 * (v[0] * tb[1] - v[1] * tb[0])};
    for (int k = 0; k < 3; k++) qb[k] += s * ub[k];
    qb[3] += ob[0] * t[0] + ob[1] * t[1] + ob[2] * t[2];
  }
  if (vb) {
    vb[0] += ob[0] + 2 * (tb[1] * u[2] - tb[2] * u[1]);
    vb[1] += ob[1] + 2 * (tb[2] * u[0] - tb[0] * u[2]);
    vb[2] += ob[2] + 2 * (tb[0] * u[1] - tb[1] * u[0]);
  }
}

// The reverse of separations (not known): from the separations' adjoints into the
// bodies' positions, starts and turns, the row's anchors, adjusted and normal. Everything it
// reads comes first (the bodies, the row's anchors, the adjoints it adds to), then the sums in
// the order the reference's terms take, then the stores: no read waits on a store.
template <int N>
HD void separations_vjp(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, RowAcc<N, true>& A,
                        int ga, int gb, bool ma, bool mb, int i, const double (&sepb)[N]) {
  const int i3 = 3 * i;
  double n[3], pa[3], pb[3], sa[3], sb[3], qa[4], qb[4];
  for (int c = 0; c < 3; c++) {
    n[c] = W.normal()[i3 + c];
    pa[c] = B.p[3 * ga + c]; pb[c] = B.p[3 * gb + c];
    sa[c] = B.start[3 * ga + c]; sb[c] = B.start[3 * gb + c];
  }
  for (int c = 0; c < 4; c++) { qa[c] = B.q[4 * ga + c]; qb[c] = B.q[4 * gb + c]; }
  double wla[N][3], wlb[N][3], lab[N][3], lbb[N][3], adj[N];
  for (int k = 0; k < N; k++) {
    const int j = P * i + k, j3 = 3 * j;
    adj[k] = WB.adjusted()[j];
    for (int c = 0; c < 3; c++) {
      wla[k][c] = W.la()[j3 + c]; wlb[k][c] = W.lb()[j3 + c];
      lab[k][c] = WB.la()[j3 + c]; lbb[k][c] = WB.lb()[j3 + c];
    }
  }
  double pab[3], sab[3], qab[4], pbb[3], sbb[3], qbb[4];
  for (int c = 0; c < 3; c++) {
    pab[c] = BB.p[3 * ga + c]; sab[c] = BB.start[3 * ga + c];
    pbb[c] = BB.p[3 * gb + c]; sbb[c] = BB.start[3 * gb + c];
  }
  for (int c = 0; c < 4; c++) { qab[c] = BB.q[4 * ga + c]; qbb[c] = BB.q[4 * gb + c]; }
  double* nb = A.normal;
  const V3 na = rotate(qa, n[0], n[1], n[2], true);
  const V3 nbv = rotate(qb, n[0], n[1], n[2], true);
  double dnb = 0, nab[3] = {0, 0, 0}, nbb[3] = {0, 0, 0};
  for (int k = 0; k < N; k++) {
    const double g = sepb[k];
    if (g == 0) continue;
    dnb += g;
    adj[k] += g;
    lbb[k][0] += g * nbv.x; lbb[k][1] += g * nbv.y; lbb[k][2] += g * nbv.z;
    lab[k][0] -= g * na.x; lab[k][1] -= g * na.y; lab[k][2] -= g * na.z;
    for (int c = 0; c < 3; c++) { nbb[c] += g * wlb[k][c]; nab[c] -= g * wla[k][c]; }
  }
  for (int k = 0; k < N; k++) {
    if (sepb[k] == 0) continue;
    const int j = P * i + k, j3 = 3 * j;
    WB.adjusted()[j] = adj[k];
    for (int c = 0; c < 3; c++) { WB.la()[j3 + c] = lab[k][c]; WB.lb()[j3 + c] = lbb[k][c]; }
  }
  for (int c = 0; c < 3; c++) {
    const double d = (pb[c] - sb[c]) - (pa[c] - sa[c]);
    nb[c] += dnb * d;
  }
  if (mb) {
    for (int c = 0; c < 3; c++) { pbb[c] += dnb * n[c]; sbb[c] -= dnb * n[c]; }
    rotate_vjp(qb, n, true, nbb, qbb, nb);
    for (int c = 0; c < 3; c++) { BB.p[3 * gb + c] = pbb[c]; BB.start[3 * gb + c] = sbb[c]; }
    for (int c = 0; c < 4; c++) BB.q[4 * gb + c] = qbb[c];
  } else {
    rotate_vjp(qb, n, true, nbb, nullptr, nb);
  }
  if (ma) {
    for (int c = 0; c < 3; c++) { pab[c] -= dnb * n[c]; sab[c] += dnb * n[c]; }
    rotate_vjp(qa, n, true, nab, qab, nb);
    for (int c = 0; c < 3; c++) { BB.p[3 * ga + c] = pab[c]; BB.start[3 * ga + c] = sab[c]; }
    for (int c = 0; c < 4; c++) BB.q[4 * ga + c] = qab[c];
  } else {
    rotate_vjp(qa, n, true, nab, nullptr, nb);
  }
}

// The reverse of one row of a sweep (KIND 0 warm, 1 solve, 2 relax), from the sweep's
// tape; the bodies' positions and turns are where the sweep found them. What the row adds
// to its bodies' inverse mass and inertia adjoints stays with the row (WB.ima() ...) until
// prepare's reverse hands it on, so a warm start's rows, which change no velocity
// adjoint, can all go back at once.
template <int N, int KIND>
HD void row_vjp_n(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, const Rows& R, int boff, int r,
                  int i, const Soft& soft, double h, const TapeRow& T, double* slot) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const bool ma = moves(B, ga), mb = moves(B, gb);
  Pair sb;
  for (int c = 0; c < 3; c++) {
    // read whether or not the body moves (every body has them), then kept only if it does
    const double va = BB.v[3 * ga + c], wa = BB.w[3 * ga + c], vb = BB.v[3 * gb + c], wb = BB.w[3 * gb + c];
    sb.v[c] = ma ? va : 0; sb.v[3 + c] = ma ? wa : 0;
    sb.v[6 + c] = mb ? vb : 0; sb.v[9 + c] = mb ? wb : 0;
  }
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  const double* Ia = B.inv_i + 6 * ga;
  const double* Ib = B.inv_i + 6 * gb;
  double imab = 0, imbb = 0;
  if (KIND == 0) {
    RowAcc<N, false>& A = *reinterpret_cast<RowAcc<N, false>*>(slot);
    A.load(WB, i);
    warm_vjp<N>(W, A, R, r, i, ima, imb, Ia, Ib, T, sb, imab, imbb);
    A.ima += imab;
    A.imb += imbb;
    A.store(WB, i);
  } else {
    RowAcc<N, true>& A = *reinterpret_cast<RowAcc<N, true>*>(slot);
    A.load(WB, i);
    // the row's tape read at once (the fields the reverse does not read are never loaded)
    double tl[ROW_TAPE];
    for (int f = 0; f < ROW_TAPE; f++) tl[f] = T[f];
    const TapeRow TL{tl, 1};
    double sepb[N];
    solve_vjp<N>(W, A, R, r, i, ima, imb, Ia, Ib, soft, h, KIND == 1, TL, sb, imab, imbb, sepb);
    separations_vjp<N>(B, BB, W, WB, A, ga, gb, ma, mb, i, sepb);
    A.ima += imab;
    A.imb += imbb;
    A.store(WB, i);
    if (ma) for (int c = 0; c < 3; c++) { BB.v[3 * ga + c] = sb.v[c]; BB.w[3 * ga + c] = sb.v[3 + c]; }
    if (mb) for (int c = 0; c < 3; c++) { BB.v[3 * gb + c] = sb.v[6 + c]; BB.w[3 * gb + c] = sb.v[9 + c]; }
  }
}

template <int KIND>
HD void row_vjp(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, const Rows& R, int boff, int r,
                int i, const Soft& soft, double h, const TapeRow& T, double* slot) {
  switch (R.count[r]) {
    case 1: row_vjp_n<1, KIND>(B, BB, W, WB, R, boff, r, i, soft, h, T, slot); break;
    case 2: row_vjp_n<2, KIND>(B, BB, W, WB, R, boff, r, i, soft, h, T, slot); break;
    case 3: row_vjp_n<3, KIND>(B, BB, W, WB, R, boff, r, i, soft, h, T, slot); break;
    default: row_vjp_n<4, KIND>(B, BB, W, WB, R, boff, r, i, soft, h, T, slot); break;
  }
}

// A sweep backwards: the overflow in reverse order, then the colours from the last.
template <int KIND, class Sync>
HD void sweep_vjp(const Ctx& c, Sync sync, const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB,
                  const Rows& R, const StepView& v, const Soft& soft, double h, const Tape& tape, int tsweep) {
  const int from = v.starts[OVERFLOW], to = v.starts[OVERFLOW + 1];
  if (from < to) {
    if (c.tid == 0)
      for (int lm = to - 1; lm >= from; lm--) {
        const int i = v.w0 + lm;
        row_vjp<KIND>(B, BB, W, WB, R, v.boff, v.r0 + lm, i, soft, h, tape.row(tsweep, i), c.slot);
      }
    sync();
  }
  for (int col = COLOURS - 1; col >= 0; col--) {
    const int f0 = v.starts[col], t0 = v.starts[col + 1];
    if (f0 == t0) continue;
    for (int lm = f0 + c.tid; lm < t0; lm += c.threads) {
      const int i = v.w0 + lm;
      row_vjp<KIND>(B, BB, W, WB, R, v.boff, v.r0 + lm, i, soft, h, tape.row(tsweep, i), c.slot);
    }
    sync();
  }
}


// The reverse of accelerate_body; B holds the body's state before it.
HD void accelerate_vjp(const Bodies& B, const BodiesBar& BB, int g, double h, int substeps) {
  double* vb = BB.v + 3 * g;
  double* wb = BB.w + 3 * g;
  if (B.capped[g]) {
    double w[3] = {B.w[3 * g], B.w[3 * g + 1], B.w[3 * g + 2]};
    if (B.angular_damping[g] != 0) {
      const double k = 1 / (1 + h * B.angular_damping[g]);
      for (int c = 0; c < 3; c++) w[c] *= k;
    }
    const double most = SPIN_TURN / (h * substeps);
    const double w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    if (w2 > most * most) {
      const double l = dsqrt(w2), k = most / l;
      const double u[3] = {w[0] / l, w[1] / l, w[2] / l};
      const double ou = wb[0] * u[0] + wb[1] * u[1] + wb[2] * u[2];
      for (int c = 0; c < 3; c++) wb[c] = k * (wb[c] - u[c] * ou);
    }
  }
  if (B.damping[g] != 0) {
    const double k = 1 / (1 + h * B.damping[g]);
    for (int c = 0; c < 3; c++) vb[c] *= k;
  }
  if (B.angular_damping[g] != 0) {
    const double k = 1 / (1 + h * B.angular_damping[g]);
    for (int c = 0; c < 3; c++) wb[c] *= k;
  }
}

// The reverse of advance_t(q, w, h): out = n / |n|, n = q + h (w, 0) q. Into qb and wb
// (overwritten) from out's adjoint ob.
HD void advance_vjp(const double* q, const double* w, double h, const double* ob, double* qb, double* wb) {
  const double nx = q[0] + h * (w[0] * q[3] + w[1] * q[2] - w[2] * q[1]);
  const double ny = q[1] + h * (w[1] * q[3] + w[2] * q[0] - w[0] * q[2]);
  const double nz = q[2] + h * (w[2] * q[3] + w[0] * q[1] - w[1] * q[0]);
  const double nw = q[3] - h * (w[0] * q[0] + w[1] * q[1] + w[2] * q[2]);
  const double l = dsqrt(nx * nx + ny * ny + nz * nz + nw * nw), il = 1 / l;
  const double ox = nx * il, oy = ny * il, oz = nz * il, ow = nw * il;
  const double oo = ob[0] * ox + ob[1] * oy + ob[2] * oz + ob[3] * ow;
  const double bx = (ob[0] - ox * oo) * il, by = (ob[1] - oy * oo) * il, bz = (ob[2] - oz * oo) * il,
               bw = (ob[3] - ow * oo) * il;
  qb[0] = bx + h * (w[2] * by - w[1] * bz - w[0] * bw);
  qb[1] = by + h * (-w[2] * bx + w[0] * bz - w[1] * bw);
  qb[2] = bz + h * (w[1] * bx - w[0] * by - w[2] * bw);
  qb[3] = bw + h * (w[0] * bx + w[1] * by + w[2] * bz);
  wb[0] = h * (q[3] * bx - q[2] * by + q[1] * bz - q[0] * bw);
  wb[1] = h * (q[2] * bx + q[3] * by - q[0] * bz - q[1] * bw);
  wb[2] = h * (-q[1] * bx + q[0] * by + q[3] * bz - q[2] * bw);
}

// The reverse of move_body; B holds the body's state before it.
HD void move_vjp(const Bodies& B, const BodiesBar& BB, int g, double h) {
  for (int c = 0; c < 3; c++) BB.v[3 * g + c] += h * BB.p[3 * g + c];
  const double* w = B.w + 3 * g;
  const bool turning = w[0] != 0 || w[1] != 0 || w[2] != 0;
  if (B.tumbles[g]) {
    if (!turning) return;
    // in: q 4, w 3, left 3, ceiling 1, inertia 3, inverse 3; out: q 4, w 3, ceiling 1 (left = w)
    double in[17], outb[8], inb[17];
    for (int k = 0; k < 4; k++) in[k] = B.q[4 * g + k];
    for (int k = 0; k < 3; k++) {
      in[4 + k] = w[k]; in[7 + k] = B.left[3 * g + k];
      in[11 + k] = B.inertia[3 * g + k]; in[14 + k] = B.inv_inertia[3 * g + k];
    }
    in[10] = B.ceiling[g];
    for (int k = 0; k < 4; k++) outb[k] = BB.q[4 * g + k];
    for (int k = 0; k < 3; k++) outb[4 + k] = BB.w[3 * g + k] + BB.left[3 * g + k];
    outb[7] = BB.ceiling[g];
    for (int k = 0; k < 17; k++) inb[k] = 0;
    vjp<DW>(in, outb, inb, [&](const Dual<DW>* x, Dual<DW>* y) {
      Dual<DW> c;
      tumble_t(x, x + 4, x + 7, x[10], x + 11, x + 14, h, y, y + 4, c);
      y[7] = c;
    });
    for (int k = 0; k < 4; k++) BB.q[4 * g + k] = inb[k];
    for (int k = 0; k < 3; k++) {
      BB.w[3 * g + k] = inb[4 + k]; BB.left[3 * g + k] = inb[7 + k];
      BB.inertia[3 * g + k] += inb[11 + k]; BB.inv_inertia[3 * g + k] += inb[14 + k];
    }
    BB.ceiling[g] = inb[10];
  } else {
    // advance (also where w is 0: the forward's skip is advance's value there)
    double qb[4], wb[3];
    advance_vjp(B.q + 4 * g, w, h * 0.5, BB.q + 4 * g, qb, wb);
    for (int k = 0; k < 4; k++) BB.q[4 * g + k] = qb[k];
    for (int k = 0; k < 3; k++) BB.w[3 * g + k] += wb[k];
  }
}

// The reverse of bounce_row, its points from the last.
HD void bounce_vjp(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, const Rows& R, int boff,
                   int r, int i, double gravity) {
  const double e = W.e()[i];
  if (e == 0) return;
  const double threshold = dsqrt(2 * gravity * R.tolerance[r]);
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const bool ma = moves(B, ga), mb = moves(B, gb);
  const double* n = W.normal() + 3 * i;
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  for (int k = R.count[r] - 1; k >= 0; k--) {
    const int j = P * i + k;
    if (W.relvel()[j] > -threshold || W.most()[j] == 0) continue;
    const double* o = W.bounce() + BOUNCE * j;
    const double old = o[0], vn = o[1], rv = W.relvel()[j], nm = W.nmass()[j];
    Pair s;
    for (int c = 0; c < 12; c++) s.v[c] = o[2 + c];
    double raw = old - nm * (vn + e * rv);
    const double next = raw < 0 ? 0 : raw;
    Pair sb;
    for (int c = 0; c < 3; c++) {
      sb.v[c] = ma ? BB.v[3 * ga + c] : 0; sb.v[3 + c] = ma ? BB.w[3 * ga + c] : 0;
      sb.v[6 + c] = mb ? BB.v[3 * gb + c] : 0; sb.v[9 + c] = mb ? BB.w[3 * gb + c] : 0;
    }
    double imab = 0, imbb = 0;
    const double lb = push_normal_vjp(sb, ima, imb, n, W.jac() + J * j, next - old, WB.normal() + 3 * i, imab, imbb,
                                      WB.jac() + J * j);
    const double nextb = WB.impulse()[j] + lb;
    double oldb = -lb;
    if (raw > 0) {
      oldb += nextb;
      WB.nmass()[j] -= nextb * (vn + e * rv);
      const double vnb = -nm * nextb;
      WB.e()[i] -= nm * nextb * rv;
      WB.relvel()[j] -= nm * e * nextb;
      closing_vjp(s, n, W.jac() + J * j, vnb, sb, WB.normal() + 3 * i, WB.jac() + J * j);
    }
    WB.impulse()[j] = oldb;
    if (ma) {
      for (int c = 0; c < 3; c++) { BB.v[3 * ga + c] = sb.v[c]; BB.w[3 * ga + c] = sb.v[3 + c]; }
      BB.inv_mass[ga] += imab;
    }
    if (mb) {
      for (int c = 0; c < 3; c++) { BB.v[3 * gb + c] = sb.v[6 + c]; BB.w[3 * gb + c] = sb.v[9 + c]; }
      BB.inv_mass[gb] += imbb;
    }
  }
}

// c = a x b: adds a's and b's adjoints from c's.
HD void cross_vjp(const double* a, const double* b, const double* cb, double* ab, double* bb) {
  if (ab) {
    ab[0] += b[1] * cb[2] - b[2] * cb[1];
    ab[1] += b[2] * cb[0] - b[0] * cb[2];
    ab[2] += b[0] * cb[1] - b[1] * cb[0];
  }
  if (bb) {
    bb[0] += cb[1] * a[2] - cb[2] * a[1];
    bb[1] += cb[2] * a[0] - cb[0] * a[2];
    bb[2] += cb[0] * a[1] - cb[1] * a[0];
  }
}

// The adjoints of turned(p, start, q, start_q) from its output's.
HD void turned_vjp(const Bodies& B, const BodiesBar& BB, int g, const double* Tb) {
  double in[14], outb[6], inb[14];
  for (int k = 0; k < 3; k++) { in[k] = B.p[3 * g + k]; in[3 + k] = B.start[3 * g + k]; }
  for (int k = 0; k < 4; k++) { in[6 + k] = B.q[4 * g + k]; in[10 + k] = B.start_q[4 * g + k]; }
  for (int k = 0; k < 6; k++) outb[k] = Tb[k];
  for (int k = 0; k < 14; k++) inb[k] = 0;
  vjp<DW>(in, outb, inb, [&](const Dual<DW>* x, Dual<DW>* y) { turned_t(x, x + 3, x + 6, x + 10, y); });
  for (int k = 0; k < 3; k++) { BB.p[3 * g + k] += inb[k]; BB.start[3 * g + k] += inb[3 + k]; }
  for (int k = 0; k < 4; k++) { BB.q[4 * g + k] += inb[6 + k]; BB.start_q[4 * g + k] += inb[10 + k]; }
}

// The reverse of store_row: from the stored impulses' adjoints (what later steps' carries
// gave them) into the row's final state, tangents and normal, and through the slip's
// turn into the bodies' poses.
HD void store_vjp(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, const Rows& R,
                  const RowsBar& RB, int boff, int r, int i) {
  const int i3 = 3 * i, count = R.count[r];
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const double* n = W.normal() + i3;
  const double* t1 = W.t1() + i3;
  const double* t2 = W.t2() + i3;
  double* nb = WB.normal() + i3;
  double* t1b = WB.t1() + i3;
  double* t2b = WB.t2() + i3;
  const double* Lb = RB.store_twist + 3 * r;
  const double* Fb = RB.store_friction + 3 * r;
  const double* Sb = RB.store_slip + 3 * r;
  for (int k = 0; k < count; k++) WB.impulse()[P * i + k] += RB.store_impulse[P * r + k];
  const double r1 = W.roll()[2 * i], r2 = W.roll()[2 * i + 1], sp = W.spin()[i];
  const double l1 = W.tangent()[2 * i], l2 = W.tangent()[2 * i + 1];
  const double d1 = W.drift()[2 * i], d2 = W.drift()[2 * i + 1];
  double d1b = 0, d2b = 0;
  for (int c = 0; c < 3; c++) {
    WB.roll()[2 * i] += t1[c] * Lb[c]; WB.roll()[2 * i + 1] += t2[c] * Lb[c]; WB.spin()[i] += n[c] * Lb[c];
    WB.tangent()[2 * i] += t1[c] * Fb[c]; WB.tangent()[2 * i + 1] += t2[c] * Fb[c];
    d1b += t1[c] * Sb[c]; d2b += t2[c] * Sb[c];
    t1b[c] += r1 * Lb[c] + l1 * Fb[c] + d1 * Sb[c];
    t2b[c] += r2 * Lb[c] + l2 * Fb[c] + d2 * Sb[c];
    nb[c] += sp * Lb[c];
  }
  double pressed = 0;
  for (int k = 0; k < count; k++) pressed += W.impulse()[P * i + k];
  if (W.sliding[i] != 0 || pressed == 0 || impacted(R, r)) {
    WB.drift()[2 * i] = 0; WB.drift()[2 * i + 1] = 0;
    return;
  }
  // drift_out = drift_in + (dx . t1, dx . t2)
  double dx[3] = {0, 0, 0};
  double T[6];
  if (B.type[gb] == DYNAMIC) {
    turned_t<double>(B.p + 3 * gb, B.start + 3 * gb, B.q + 4 * gb, B.start_q + 4 * gb, T);
    const double* rr = W.anchor_b() + i3;
    dx[0] += T[0] + T[4] * rr[2] - T[5] * rr[1];
    dx[1] += T[1] + T[5] * rr[0] - T[3] * rr[2];
    dx[2] += T[2] + T[3] * rr[1] - T[4] * rr[0];
  }
  if (B.type[ga] == DYNAMIC) {
    turned_t<double>(B.p + 3 * ga, B.start + 3 * ga, B.q + 4 * ga, B.start_q + 4 * ga, T);
    const double* rr = W.anchor_a() + i3;
    dx[0] -= T[0] + T[4] * rr[2] - T[5] * rr[1];
    dx[1] -= T[1] + T[5] * rr[0] - T[3] * rr[2];
    dx[2] -= T[2] + T[3] * rr[1] - T[4] * rr[0];
  }
  WB.drift()[2 * i] = d1b; WB.drift()[2 * i + 1] = d2b;
  double dxb[3];
  for (int c = 0; c < 3; c++) {
    dxb[c] = d1b * t1[c] + d2b * t2[c];
    t1b[c] += d1b * dx[c];
    t2b[c] += d2b * dx[c];
  }
  for (int side = 0; side < 2; side++) {
    const int g = side == 0 ? gb : ga;
    if (B.type[g] != DYNAMIC) continue;
    const double sg = side == 0 ? 1 : -1;
    turned_t<double>(B.p + 3 * g, B.start + 3 * g, B.q + 4 * g, B.start_q + 4 * g, T);
    const double* rr = (side == 0 ? W.anchor_b() : W.anchor_a()) + i3;
    double* rrb = (side == 0 ? WB.anchor_b() : WB.anchor_a()) + i3;
    const double cb[3] = {sg * dxb[0], sg * dxb[1], sg * dxb[2]};
    double Tb[6] = {cb[0], cb[1], cb[2], 0, 0, 0};
    cross_vjp(T + 3, rr, cb, Tb + 3, rrb);
    turned_vjp(B, BB, g, Tb);
  }
}

// The reverse of lever(): from the adjoints of jac (12) and of the returned mass.
HD void lever_vjp(const double* Ia, const double* Ib, double ima, double imb, const double* ra, const double* rb,
                  const double* d, const double* jacb, double massb, double* rab, double* rbb, double* db, double* Iab,
                  double* Ibb, double& imab, double& imbb) {
  double ca[3], cb[3], ia[3], ib[3];
  ca[0] = ra[1] * d[2] - ra[2] * d[1]; ca[1] = ra[2] * d[0] - ra[0] * d[2]; ca[2] = ra[0] * d[1] - ra[1] * d[0];
  cb[0] = rb[1] * d[2] - rb[2] * d[1]; cb[1] = rb[2] * d[0] - rb[0] * d[2]; cb[2] = rb[0] * d[1] - rb[1] * d[0];
  ia[0] = Ia[0] * ca[0] + Ia[1] * ca[1] + Ia[2] * ca[2];
  ia[1] = Ia[1] * ca[0] + Ia[3] * ca[1] + Ia[4] * ca[2];
  ia[2] = Ia[2] * ca[0] + Ia[4] * ca[1] + Ia[5] * ca[2];
  ib[0] = Ib[0] * cb[0] + Ib[1] * cb[1] + Ib[2] * cb[2];
  ib[1] = Ib[1] * cb[0] + Ib[3] * cb[1] + Ib[4] * cb[2];
  ib[2] = Ib[2] * cb[0] + Ib[4] * cb[1] + Ib[5] * cb[2];
  const double k = ima + imb + ca[0] * ia[0] + ca[1] * ia[1] + ca[2] * ia[2] + cb[0] * ib[0] + cb[1] * ib[1] + cb[2] * ib[2];
  const double kb = k > 0 ? -massb / (k * k) : 0;
  imab += kb; imbb += kb;
  double iab[3], ibb[3], cab[3], cbb[3];
  for (int c = 0; c < 3; c++) {
    iab[c] = jacb[6 + c] + kb * ca[c];
    ibb[c] = jacb[9 + c] + kb * cb[c];
    cab[c] = jacb[c] + kb * ia[c];
    cbb[c] = jacb[3 + c] + kb * ib[c];
  }
  sym_vjp(Ia, ca, iab, Iab, cab);
  sym_vjp(Ib, cb, ibb, Ibb, cbb);
  cross_vjp(ra, d, cab, rab, db);
  cross_vjp(rb, d, cbb, rbb, db);
}

// The reverse of turning(Ia, Ib, x): 1 / (x^T (Ia + Ib) x), 0 if that is not positive.
HD void turning_vjp(const double* Ia, const double* Ib, const double* x, double mb, double* Iab, double* Ibb,
                    double* xb) {
  const double k = along(Ia, x[0], x[1], x[2]) + along(Ib, x[0], x[1], x[2]);
  if (!(k > 0) || mb == 0) return;
  const double kb = -mb / (k * k);
  // d along / dI: x x^T on the stored entries; d along / dx = 2 I x
  const double o[6] = {x[0] * x[0], 2 * x[0] * x[1], 2 * x[0] * x[2], x[1] * x[1], 2 * x[1] * x[2], x[2] * x[2]};
  for (int c = 0; c < 6; c++) {
    if (Iab) Iab[c] += kb * o[c];
    if (Ibb) Ibb[c] += kb * o[c];
  }
  for (int c = 0; c < 3; c++) {
    const double s = 2 * kb;
    xb[c] += s * ((Ia[c == 0 ? 0 : c == 1 ? 1 : 2] * x[0] + Ia[c == 0 ? 1 : c == 1 ? 3 : 4] * x[1]
                   + Ia[c == 0 ? 2 : c == 1 ? 4 : 5] * x[2])
                  + (Ib[c == 0 ? 0 : c == 1 ? 1 : 2] * x[0] + Ib[c == 0 ? 1 : c == 1 ? 3 : 4] * x[1]
                     + Ib[c == 0 ? 2 : c == 1 ? 4 : 5] * x[2]));
  }
}

// The reverse of core(): from the adjoints of the anchor in the body's frame (lb) and
// of the returned world offset (cb), into r and the body's turn at the start.
HD void core_vjp(const Bodies& B, const BodiesBar& BB, const Sizes& Z, int g, const double* r, const double* lb,
                 const double* cb, double* rb) {
  const int shape = B.shape[g];
  if (shape == SPHERE_SHAPE) return;
  const double* q = B.q + 4 * g;
  double* qb = moves(B, g) ? BB.q + 4 * g : nullptr;
  if (shape != CAPSULE_SHAPE) {
    for (int c = 0; c < 3; c++) rb[c] += cb[c];
    rotate_vjp(q, r, true, lb, qb, rb);
    return;
  }
  const double half = Z.segment[g];
  const V3 v = rotate(q, r[0], r[1], r[2], true);
  const double cc = v.y < -half ? -half : v.y > half ? half : v.y;
  double ccb = lb[1];
  const double e[3] = {0, cc, 0};
  double eb[3] = {0, 0, 0};
  rotate_vjp(q, e, false, cb, qb, eb);
  ccb += eb[1];
  if (v.y < -half || v.y > half) return;
  const double vb[3] = {0, ccb, 0};
  rotate_vjp(q, r, true, vb, qb, rb);
}

// The reverse of prepare_row; the bodies are at the step's start.
HD void prepare_vjp(const Bodies& B, const BodiesBar& BB, const Sizes& Z, const Work& W, const WorkBar& WB,
                    const Rows& R, int boff, int r, int i) {
  const int ga = boff + R.a[r], gb = boff + R.b[r];
  const bool ma = moves(B, ga), mb = moves(B, gb);
  const double* Ia = B.inv_i + 6 * ga;
  const double* Ib = B.inv_i + 6 * gb;
  double Iab[6] = {0, 0, 0, 0, 0, 0}, Ibb[6] = {0, 0, 0, 0, 0, 0};
  double imab = 0, imbb = 0;
  const double *pa = B.p + 3 * ga, *pb = B.p + 3 * gb;
  const double *va = B.v + 3 * ga, *vb = B.v + 3 * gb;
  const double *wa = B.w + 3 * ga, *wb = B.w + 3 * gb;
  const double ima = B.inv_mass[ga], imb = B.inv_mass[gb];
  const int i3 = 3 * i, count = R.count[r];
  const double* n = W.normal() + i3;
  const double* t1 = W.t1() + i3;
  const double* t2 = W.t2() + i3;
  double* nb = WB.normal() + i3;
  double* t1b = WB.t1() + i3;
  double* t2b = WB.t2() + i3;
  const double tol = R.tolerance[r];
  double ra[P][3], rb[P][3], rab[P][3], rbb[P][3], sepb[P];
  double near[P], weight = 0;
  for (int k = 0; k < count; k++) {
    const int j3 = 3 * (P * i + k);
    for (int c = 0; c < 3; c++) {
      ra[k][c] = W.point()[j3 + c] - pa[c];
      rb[k][c] = W.point()[j3 + c] + R.shift[3 * (P * r + k) + c] - pb[c];
      rab[k][c] = 0; rbb[k][c] = 0;
    }
    sepb[k] = 0;
    const double s = W.separation()[P * i + k] / tol;
    near[k] = s <= 1 ? 1 : s >= 2 ? 0 : 2 - s;
    weight += near[k];
  }
  const double* A = W.anchor_a() + i3;
  const double* Bn = W.anchor_b() + i3;
  double Ab[3] = {WB.anchor_a()[i3], WB.anchor_a()[i3 + 1], WB.anchor_a()[i3 + 2]};
  double Bb[3] = {WB.anchor_b()[i3], WB.anchor_b()[i3 + 1], WB.anchor_b()[i3 + 2]};
  // the tangent masses from the friction lever terms
  {
    const double* f = W.fjac() + FJ * i;
    const double* g = f + 12;
    double fb[FJ];
    for (int c = 0; c < FJ; c++) fb[c] = WB.fjac()[FJ * i + c];
    const double k11 = ima + imb + f[0] * f[6] + f[1] * f[7] + f[2] * f[8] + f[3] * f[9] + f[4] * f[10] + f[5] * f[11];
    const double k22 = ima + imb + g[0] * g[6] + g[1] * g[7] + g[2] * g[8] + g[3] * g[9] + g[4] * g[10] + g[5] * g[11];
    const double k12 = g[0] * f[6] + g[1] * f[7] + g[2] * f[8] + g[3] * f[9] + g[4] * f[10] + g[5] * f[11];
    const double det = k11 * k22 - k12 * k12;
    if (det > 0) {
      const double* tmb = WB.tm() + i3;
      const double detb = (-k22 * tmb[0] + k12 * tmb[1] - k11 * tmb[2]) / (det * det);
      const double k11b = tmb[2] / det + detb * k22;
      const double k22b = tmb[0] / det + detb * k11;
      const double k12b = -tmb[1] / det - 2 * k12 * detb;
      imab += k11b + k22b; imbb += k11b + k22b;
      for (int c = 0; c < 6; c++) {
        fb[c] += k11b * f[6 + c]; fb[6 + c] += k11b * f[c];
        fb[12 + c] += k22b * g[6 + c]; fb[18 + c] += k22b * g[c];
        fb[12 + c] += k12b * f[6 + c]; fb[6 + c] += k12b * g[c];
      }
    }
    lever_vjp(Ia, Ib, ima, imb, A, Bn, t1, fb, 0, Ab, Bb, t1b, Iab, Ibb, imab, imbb);
    lever_vjp(Ia, Ib, ima, imb, A, Bn, t2, fb + 12, 0, Ab, Bb, t2b, Iab, Ibb, imab, imbb);
  }
  // the twist levers: |e - (e . n) n|, e = ra - anchor a
  for (int k = 0; k < count; k++) {
    const double lb = WB.lever()[P * i + k];
    if (lb == 0) continue;
    double e[3] = {ra[k][0] - A[0], ra[k][1] - A[1], ra[k][2] - A[2]};
    const double al = e[0] * n[0] + e[1] * n[1] + e[2] * n[2];
    double d[3] = {e[0] - al * n[0], e[1] - al * n[1], e[2] - al * n[2]};
    const double l = dsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(l > 0)) continue;
    double dbv[3] = {lb * d[0] / l, lb * d[1] / l, lb * d[2] / l};
    const double nd = n[0] * dbv[0] + n[1] * dbv[1] + n[2] * dbv[2];
    for (int c = 0; c < 3; c++) {
      const double eb = dbv[c] - n[c] * nd;
      rab[k][c] += eb; Ab[c] -= eb;
      nb[c] += -al * dbv[c] - nd * e[c];
    }
  }
  // the anchors
  if (weight == 0) {
    for (int k = 0; k < count; k++)
      for (int c = 0; c < 3; c++) { rab[k][c] += Ab[c] / count; rbb[k][c] += Bb[c] / count; }
  } else {
    for (int k = 0; k < count; k++) {
      double nearb = 0;
      for (int c = 0; c < 3; c++) {
        rab[k][c] += near[k] / weight * Ab[c];
        rbb[k][c] += near[k] / weight * Bb[c];
        nearb += ((ra[k][c] - A[c]) * Ab[c] + (rb[k][c] - Bn[c]) * Bb[c]) / weight;
      }
      const double s = W.separation()[P * i + k] / tol;
      if (s > 1 && s < 2) sepb[k] -= nearb / tol;
    }
  }
  // the block: K[k][l] = ima + imb + jac_k[0..5] . jac_l[6..11]
  if (count >= 2) {
    const double* Kb = WB.block() + P * P * i;
    for (int k = 0; k < count; k++) {
      const double* jk = W.jac() + J * (P * i + k);
      double* jkb = WB.jac() + J * (P * i + k);
      for (int l = 0; l <= k; l++) {
        const double* jl = W.jac() + J * (P * i + l);
        double* jlb = WB.jac() + J * (P * i + l);
        const double eb = k == l ? Kb[P * k + k] : Kb[P * k + l] + Kb[P * l + k];
        if (eb == 0) continue;
        imab += eb; imbb += eb;
        for (int c = 0; c < 6; c++) { jkb[c] += eb * jl[6 + c]; jlb[6 + c] += eb * jk[c]; }
      }
    }
  }
  // each point: closing speed at the start, lever terms and mass, adjusted, anchors in
  // the bodies' frames
  for (int k = 0; k < count; k++) {
    const int j = P * i + k, j3 = 3 * j;
    {
      const double g = WB.relvel()[j];
      if (g != 0) {
        double dv[3];
        dv[0] = vb[0] + wb[1] * rb[k][2] - wb[2] * rb[k][1] - va[0] - wa[1] * ra[k][2] + wa[2] * ra[k][1];
        dv[1] = vb[1] + wb[2] * rb[k][0] - wb[0] * rb[k][2] - va[1] - wa[2] * ra[k][0] + wa[0] * ra[k][2];
        dv[2] = vb[2] + wb[0] * rb[k][1] - wb[1] * rb[k][0] - va[2] - wa[0] * ra[k][1] + wa[1] * ra[k][0];
        const double dvb[3] = {g * n[0], g * n[1], g * n[2]};
        const double mdvb[3] = {-dvb[0], -dvb[1], -dvb[2]};
        for (int c = 0; c < 3; c++) nb[c] += g * dv[c];
        if (mb) for (int c = 0; c < 3; c++) BB.v[3 * gb + c] += dvb[c];
        if (ma) for (int c = 0; c < 3; c++) BB.v[3 * ga + c] -= dvb[c];
        cross_vjp(wb, rb[k], dvb, mb ? BB.w + 3 * gb : nullptr, rbb[k]);
        cross_vjp(wa, ra[k], mdvb, ma ? BB.w + 3 * ga : nullptr, rab[k]);
      }
    }
    lever_vjp(Ia, Ib, ima, imb, ra[k], rb[k], n, WB.jac() + J * j, WB.nmass()[j], rab[k], rbb[k], nb, Iab, Ibb, imab, imbb);
    {
      // adjusted = separation - ((cb - ca) . n)
      const double g = WB.adjusted()[j];
      double la_[3], lb_[3];
      const V3 ca = core(B, Z, ga, ra[k][0], ra[k][1], ra[k][2], la_);
      const V3 cbv = core(B, Z, gb, rb[k][0], rb[k][1], rb[k][2], lb_);
      sepb[k] += g;
      nb[0] -= g * (cbv.x - ca.x); nb[1] -= g * (cbv.y - ca.y); nb[2] -= g * (cbv.z - ca.z);
      const double cab[3] = {g * n[0], g * n[1], g * n[2]};
      const double cbb[3] = {-g * n[0], -g * n[1], -g * n[2]};
      core_vjp(B, BB, Z, ga, ra[k], WB.la() + j3, cab, rab[k]);
      core_vjp(B, BB, Z, gb, rb[k], WB.lb() + j3, cbb, rbb[k]);
    }
    for (int c = 0; c < 3; c++) {
      WB.point()[j3 + c] += rab[k][c] + rbb[k][c];
      if (ma) BB.p[3 * ga + c] -= rab[k][c];
      if (mb) BB.p[3 * gb + c] -= rbb[k][c];
    }
    WB.separation()[j] += sepb[k];
  }
  if (count <= 2) {
    turning_vjp(Ia, Ib, t1, WB.rm()[2 * i], Iab, Ibb, t1b);
    turning_vjp(Ia, Ib, t2, WB.rm()[2 * i + 1], Iab, Ibb, t2b);
  }
  turning_vjp(Ia, Ib, n, WB.sm()[i], Iab, Ibb, nb);
  if (ma) {
    for (int c = 0; c < 6; c++) BB.inv_i[6 * ga + c] += Iab[c] + WB.Ia()[6 * i + c];
    BB.inv_mass[ga] += imab + WB.ima()[i];
  }
  if (mb) {
    for (int c = 0; c < 6; c++) BB.inv_i[6 * gb + c] += Ibb[c] + WB.Ib()[6 * i + c];
    BB.inv_mass[gb] += imbb + WB.imb()[i];
  }
}

// The reverse of place_row and carry_row: the places' adjoints into the bodies' poses at
// the step's start; the carried impulses' into the rows that stored them.
HD void place_vjp(const Bodies& B, const BodiesBar& BB, const Work& W, const WorkBar& WB, const Rows& R,
                  const RowsBar& RB, int boff, int r, int i) {
  const int ga = boff + R.a[r], gb = boff + R.b[r], count = R.count[r], i3 = 3 * i;
  const double* t1 = W.t1() + i3;
  const double* t2 = W.t2() + i3;
  const double* n = W.normal() + i3;
  // carry
  const int s = R.src[r];
  if (s >= 0) {
    const double* L = R.store_twist + 3 * s;
    const double* F = R.store_friction + 3 * s;
    double* Lb = RB.store_twist + 3 * s;
    double* Fb = RB.store_friction + 3 * s;
    const double r1b = count <= 2 ? WB.roll()[2 * i] : 0, r2b = count <= 2 ? WB.roll()[2 * i + 1] : 0;
    const double sb = WB.spin()[i], o1b = WB.tangent()[2 * i], o2b = WB.tangent()[2 * i + 1];
    for (int c = 0; c < 3; c++) {
      Lb[c] += r1b * t1[c] + r2b * t2[c] + sb * n[c];
      Fb[c] += o1b * t1[c] + o2b * t2[c];
      WB.t1()[i3 + c] += r1b * L[c] + o1b * F[c];
      WB.t2()[i3 + c] += r2b * L[c] + o2b * F[c];
      WB.normal()[i3 + c] += sb * L[c];
    }
    bool same = false;
    for (int k = 0; k < count; k++) {
      const uint32_t id = R.pid[P * r + k];
      for (int j = 0; j < R.count[s]; j++) {
        if (R.pid[P * s + j] != id) continue;
        RB.store_impulse[P * s + j] += WB.impulse()[P * i + k];
        same = true;
        break;
      }
    }
    if (same) {
      const double* S = R.store_slip + 3 * s;
      double* Sb = RB.store_slip + 3 * s;
      const double d1b = WB.drift()[2 * i], d2b = WB.drift()[2 * i + 1];
      for (int c = 0; c < 3; c++) {
        Sb[c] += d1b * t1[c] + d2b * t2[c];
        WB.t1()[i3 + c] += d1b * S[c];
        WB.t2()[i3 + c] += d2b * S[c];
      }
    }
  }
  RB.mu[r] += WB.mu()[i];
  RB.e[r] += WB.e()[i];
  // place
  const bool ma = moves(B, ga), mb = moves(B, gb);
  if (!ma && !mb) return;
  const Held h = held_of(B, R, boff, r);
  double in[14], outb[9 + 4 * P], inb[14];
  for (int c = 0; c < 3; c++) { in[c] = B.p[3 * ga + c]; in[7 + c] = B.p[3 * gb + c]; }
  for (int c = 0; c < 4; c++) { in[3 + c] = B.q[4 * ga + c]; in[10 + c] = B.q[4 * gb + c]; }
  for (int c = 0; c < 3; c++) { outb[c] = WB.normal()[i3 + c]; outb[3 + c] = WB.t1()[i3 + c]; outb[6 + c] = WB.t2()[i3 + c]; }
  for (int k = 0; k < P; k++) {
    for (int c = 0; c < 3; c++) outb[9 + 4 * k + c] = k < count ? WB.point()[3 * (P * i + k) + c] : 0;
    outb[9 + 4 * k + 3] = k < count ? WB.separation()[P * i + k] : 0;
  }
  for (int c = 0; c < 14; c++) inb[c] = 0;
  vjp<DW>(in, outb, inb, [&](const Dual<DW>* x, Dual<DW>* y) {
    for (int k = 0; k < 9 + 4 * P; k++) y[k] = dual<DW>(0);
    place_t(h, x, x + 3, x + 7, x + 10, y);
  });
  if (ma) {
    for (int c = 0; c < 3; c++) BB.p[3 * ga + c] += inb[c];
    for (int c = 0; c < 4; c++) BB.q[4 * ga + c] += inb[3 + c];
  }
  if (mb) {
    for (int c = 0; c < 3; c++) BB.p[3 * gb + c] += inb[7 + c];
    for (int c = 0; c < 4; c++) BB.q[4 * gb + c] += inb[10 + c];
  }
}

// The reverse of inertia_body and of the step's start copies, for a dynamic body.
HD void inertia_vjp(const Bodies& B, const BodiesBar& BB, int g) {
  double in[7], outb[6], inb[7];
  for (int k = 0; k < 4; k++) in[k] = B.q[4 * g + k];
  for (int k = 0; k < 3; k++) in[4 + k] = B.inv_inertia[3 * g + k];
  for (int k = 0; k < 6; k++) { outb[k] = BB.inv_i[6 * g + k]; BB.inv_i[6 * g + k] = 0; }
  for (int k = 0; k < 7; k++) inb[k] = 0;
  vjp<DW>(in, outb, inb, [&](const Dual<DW>* x, Dual<DW>* y) { world_inverse_t(x, x + 4, y); });
  for (int k = 0; k < 4; k++) BB.q[4 * g + k] += inb[k] + BB.start_q[4 * g + k];
  for (int k = 0; k < 3; k++) {
    BB.inv_inertia[3 * g + k] += inb[4 + k];
    BB.p[3 * g + k] += BB.start[3 * g + k];
  }
  for (int k = 0; k < 4; k++) BB.start_q[4 * g + k] = 0;
  for (int k = 0; k < 3; k++) BB.start[3 * g + k] = 0;
}

// The carried drift of row r, as carry_row left it (store_row adds the step's slip).
HD void carried_drift(const Rows& R, const Work& W, int r, int i) {
  W.drift()[2 * i] = 0; W.drift()[2 * i + 1] = 0;
  const int s = R.src[r];
  if (s < 0) return;
  bool same = false;
  for (int k = 0; k < R.count[r] && !same; k++)
    for (int j = 0; j < R.count[s]; j++)
      if (R.pid[P * s + j] == R.pid[P * r + k]) { same = true; break; }
  if (!same) return;
  const double *t1 = W.t1() + 3 * i, *t2 = W.t2() + 3 * i, *S = R.store_slip + 3 * s;
  W.drift()[2 * i] = S[0] * t1[0] + S[1] * t1[1] + S[2] * t1[2];
  W.drift()[2 * i + 1] = S[0] * t2[0] + S[1] * t2[1] + S[2] * t2[2];
}

HD void zero_row_bars(const WorkBar& WB, int i) {
  for (int k = 0; k < P; k++) {
    const int j = P * i + k;
    WB.impulse()[j] = 0; WB.relvel()[j] = 0; WB.nmass()[j] = 0; WB.lever()[j] = 0; WB.adjusted()[j] = 0;
    WB.separation()[j] = 0;
    for (int c = 0; c < 3; c++) { WB.la()[3 * j + c] = 0; WB.lb()[3 * j + c] = 0; WB.point()[3 * j + c] = 0; }
    for (int c = 0; c < J; c++) WB.jac()[J * j + c] = 0;
  }
  for (int c = 0; c < P * P; c++) WB.block()[P * P * i + c] = 0;
  for (int c = 0; c < FJ; c++) WB.fjac()[FJ * i + c] = 0;
  for (int c = 0; c < 2; c++) { WB.tangent()[2 * i + c] = 0; WB.roll()[2 * i + c] = 0; WB.drift()[2 * i + c] = 0; WB.rm()[2 * i + c] = 0; }
  for (int c = 0; c < 3; c++) {
    WB.tm()[3 * i + c] = 0; WB.normal()[3 * i + c] = 0; WB.t1()[3 * i + c] = 0; WB.t2()[3 * i + c] = 0;
    WB.anchor_a()[3 * i + c] = 0; WB.anchor_b()[3 * i + c] = 0;
  }
  WB.spin()[i] = 0; WB.sm()[i] = 0; WB.mu()[i] = 0; WB.e()[i] = 0; WB.ima()[i] = 0; WB.imb()[i] = 0;
  for (int c = 0; c < 6; c++) { WB.Ia()[6 * i + c] = 0; WB.Ib()[6 * i + c] = 0; }
}

// Rows of the step in colour order, a barrier after each colour, the overflow in order:
// no two rows of a colour share a body the step moves, so what a row adds to its bodies'
// adjoints never races.
template <class Sync, class F>
HD void by_colour(const Ctx& c, Sync sync, const StepView& v, F f) {
  for (int col = 0; col < COLOURS; col++) {
    const int from = v.starts[col], to = v.starts[col + 1];
    if (from == to) continue;
    for (int lm = from + c.tid; lm < to; lm += c.threads) f(lm);
    sync();
  }
  const int from = v.starts[OVERFLOW], to = v.starts[OVERFLOW + 1];
  if (from < to) {
    if (c.tid == 0)
      for (int lm = from; lm < to; lm++) f(lm);
    sync();
  }
}

// The reverse of step_forward for one world: the step's forward has just run with its
// tape, so B, W and the tape hold everything it found; BB holds the adjoints of the
// state after the step and leaves with those of the state before it.
template <class Sync>
HD void step_reverse(const Ctx& c, Sync sync, const Bodies& B, const BodiesBar& BB, const Sizes& Z, const Rows& R,
                     const RowsBar& RB, const Work& W, const WorkBar& WB, const StepView& v, const Tape& tape) {
  const int nrows = v.r1 - v.r0;
  const bool touching = nrows > 0;
  const int substeps = v.substeps;
  const double h = v.dt / substeps;
  for (int lm = c.tid; lm < nrows; lm += c.threads) zero_row_bars(WB, v.w0 + lm);
  sync();
  if (touching) {
    by_colour(c, sync, v, [&](int lm) { store_vjp(B, BB, W, WB, R, RB, v.boff, v.r0 + lm, v.w0 + lm); });
    for (int lm = c.tid; lm < nrows; lm += c.threads) carried_drift(R, W, v.r0 + lm, v.w0 + lm);
    if (c.tid == 0) {
      const double gravity = dsqrt(v.gx * v.gx + v.gy * v.gy + v.gz * v.gz);
      for (int f = nrows - 1; f >= 0; f--) {
        const int lm = v.narrow[f];
        bounce_vjp(B, BB, W, WB, R, v.boff, v.r0 + lm, v.w0 + lm, gravity);
      }
    }
    sync();
  }
  const Soft soft = make_soft(STIFFNESS / h, DAMPING_RATIO, h);
  for (int sub = substeps - 1; sub >= 0; sub--) {
    if (touching) sweep_vjp<2>(c, sync, B, BB, W, WB, R, v, soft, h, tape, 2 * sub + 1);
    for (int s = c.tid; s < v.nbodies; s += c.threads) {
      const int g = v.boff + s;
      if (B.type[g] != DYNAMIC) continue;
      load_state(B, g, tape.body(2 * sub + 1, g));
      move_vjp(B, BB, g, h);
    }
    sync();
    if (touching) {
      sweep_vjp<1>(c, sync, B, BB, W, WB, R, v, soft, h, tape, 2 * sub);
      // a warm start's rows leave every velocity adjoint as it is: all of them at once
      for (int lm = c.tid; lm < nrows; lm += c.threads)
        row_vjp<0>(B, BB, W, WB, R, v.boff, v.r0 + lm, v.w0 + lm, soft, h, tape.row(2 * sub, v.w0 + lm), c.slot);
      sync();
    }
    for (int s = c.tid; s < v.nbodies; s += c.threads) {
      const int g = v.boff + s;
      if (B.type[g] != DYNAMIC) continue;
      load_state(B, g, tape.body(2 * sub, g));
      accelerate_vjp(B, BB, g, h, substeps);
    }
    sync();
  }
  if (touching) {
    by_colour(c, sync, v, [&](int lm) {
      prepare_vjp(B, BB, Z, W, WB, R, v.boff, v.r0 + lm, v.w0 + lm);
      place_vjp(B, BB, W, WB, R, RB, v.boff, v.r0 + lm, v.w0 + lm);
    });
  }
  for (int s = c.tid; s < v.nbodies; s += c.threads) {
    const int g = v.boff + s;
    if (B.type[g] == DYNAMIC) inertia_vjp(B, BB, g);
  }
  sync();
}


// ---------------------------------------------------------------------------------
// A batch of held trajectories, and a world's whole forward and backward pass.

struct Trajectory {
  int worlds, steps, every, bodies, substeps_most, work_rows;
  const int* body_offset;  // worlds + 1
  const int* row_offset;   // steps * worlds + 1: the rows of step t, world w from [t * worlds + w]
  const int* starts;       // STARTS per (step, world), relative to its first row
  const int* narrow;       // per row: the rows of its (step, world) in narrow order, relative
  const int* work_at;      // per (step, world): its first working row (each world has its own block)
  const int* substeps;     // per world
  const double *dt, *gravity;
  Bodies B;
  Sizes Z;
  Rows R;
  Work W;
  BodiesBar BB;
  WorkBar WB;
  RowsBar RB;
  double* checkpoints;  // (steps / every + 1) * bodies * STATE
  double *row_tape, *body_tape;
  uint64_t *body_branch, *row_branch;  // the last forward pass's branch records (note())
};

HD StepView view_of(const Trajectory& T, int t, int wld) {
  StepView v;
  v.boff = T.body_offset[wld];
  v.nbodies = T.body_offset[wld + 1] - v.boff;
  v.r0 = T.row_offset[t * T.worlds + wld];
  v.r1 = T.row_offset[t * T.worlds + wld + 1];
  v.w0 = T.work_at[t * T.worlds + wld];
  v.substeps = T.substeps[wld];
  v.dt = T.dt[wld];
  v.gx = T.gravity[3 * wld]; v.gy = T.gravity[3 * wld + 1]; v.gz = T.gravity[3 * wld + 2];
  v.starts = T.starts + (size_t)(t * T.worlds + wld) * STARTS;
  v.narrow = T.narrow + v.r0;
  return v;
}

// The tape of the k-th step of a segment.
HD Tape tape_of(const Trajectory& T, int k) {
  Tape tp;
  tp.rows = T.row_tape;
  tp.row_stride = T.work_rows;
  tp.body_stride = T.bodies;
  tp.bodies = T.body_tape + (size_t)k * 2 * T.substeps_most * T.bodies * STATE;
  return tp;
}
HD Tape no_tape() { return Tape{nullptr, nullptr, 0, 0}; }

// The columns that follow from a dynamic body's mass: its inverse mass and inertias
// scaled by mass0 / mass (exactly the world's own at mass0).
HD void from_mass(const Bodies& B, int g) {
  if (B.type[g] != DYNAMIC) {
    B.inv_mass[g] = B.inv_mass0[g];
    for (int k = 0; k < 3; k++) { B.inertia[3 * g + k] = B.inertia0[3 * g + k]; B.inv_inertia[3 * g + k] = B.inv_inertia0[3 * g + k]; }
    return;
  }
  const double down = B.mass0[g] / B.mass[g], up = B.mass[g] / B.mass0[g];
  B.inv_mass[g] = B.inv_mass0[g] * down;
  for (int k = 0; k < 3; k++) {
    B.inertia[3 * g + k] = B.inertia0[3 * g + k] * up;
    B.inv_inertia[3 * g + k] = B.inv_inertia0[3 * g + k] * down;
  }
}

// World wld's forward pass from the state in B (its mass columns already set), keeping
// its bodies' state every `every` steps.
template <class Sync>
HD void forward_world(const Ctx& c, Sync sync, const Trajectory& T, int wld) {
  const int boff = T.body_offset[wld], nb = T.body_offset[wld + 1] - boff;
  // This pass keeps its branch records (the backward pass's steps again do not).
  Bodies B = T.B;
  Rows R = T.R;
  B.branch = T.body_branch;
  R.branch = T.row_branch;
  for (int s = c.tid; s < nb; s += c.threads) B.branch[boff + s] = 0;
  // A row's record starts from what the row is (its bodies, points, feature ids, and which
  // row of its step its pair carried from), so records of two recordings differ where
  // their contacts do.
  for (int t = 0; t < T.steps; t++) {
    const int r0 = T.row_offset[t * T.worlds + wld], r1 = T.row_offset[t * T.worlds + wld + 1];
    for (int r = r0 + c.tid; r < r1; r += c.threads) {
      uint64_t x = 0;
      note(&x, ((uint64_t)R.count[r] << 56) | ((uint64_t)(uint32_t)R.a[r] << 28) | (uint64_t)(uint32_t)R.b[r]);
      for (int k = 0; k < R.count[r]; k++) note(&x, R.pid[P * r + k]);
      note(&x, R.src[r] < 0 ? ~0ull : (uint64_t)(r - R.src[r]));
      R.branch[r] = x;
    }
  }
  sync();
  for (int t = 0; t < T.steps; t++) {
    if (t % T.every == 0) {
      double* ck = T.checkpoints + (size_t)(t / T.every) * T.bodies * STATE;
      for (int s = c.tid; s < nb; s += c.threads) save_state(B, boff + s, ck + (size_t)(boff + s) * STATE);
    }
    const StepView v = view_of(T, t, wld);
    step_forward(c, sync, B, T.Z, R, T.W, v, no_tape(), false);
  }
}

// World wld's backward pass: BB's state adjoints hold those of the final state and leave
// with those of the initial one; the mass columns' adjoints and the rows' mu and e
// adjoints are summed on the way.
template <class Sync>
HD void backward_world(const Ctx& c, Sync sync, const Trajectory& T, int wld) {
  const int boff = T.body_offset[wld], nb = T.body_offset[wld + 1] - boff;
  const int segments = (T.steps + T.every - 1) / T.every;
  for (int seg = segments - 1; seg >= 0; seg--) {
    const int t0 = seg * T.every, t1 = t0 + T.every < T.steps ? t0 + T.every : T.steps;
    const double* ck = T.checkpoints + (size_t)seg * T.bodies * STATE;
    for (int s = c.tid; s < nb; s += c.threads) load_state(T.B, boff + s, ck + (size_t)(boff + s) * STATE);
    sync();
    for (int t = t0; t < t1; t++) {
      const StepView v = view_of(T, t, wld);
      step_forward(c, sync, T.B, T.Z, T.R, T.W, v, tape_of(T, t - t0), false);
    }
    for (int t = t1 - 1; t >= t0; t--) {
      const Tape tp = tape_of(T, t - t0);
      for (int s = c.tid; s < nb; s += c.threads) {
        const int g = boff + s;
        if (T.B.type[g] != DYNAMIC) continue;
        if (t < t1 - 1) load_state(T.B, g, tape_of(T, t + 1 - t0).body(0, g));
        const double* o = tp.body(0, g);
        for (int k = 0; k < 3; k++) T.B.start[3 * g + k] = o[k];
        for (int k = 0; k < 4; k++) T.B.start_q[4 * g + k] = o[3 + k];
        world_inverse_t<double>(o + 3, T.B.inv_inertia + 3 * g, T.B.inv_i + 6 * g);
      }
      sync();
      const StepView v = view_of(T, t, wld);
      step_reverse(c, sync, T.B, T.BB, T.Z, T.R, T.RB, T.W, T.WB, v, tp);
    }
  }
}

// World wld's parameter adjoints, summed in a fixed order: each row's mu = sqrt(fa fb)
// and e = max(ea, eb) onto its bodies' friction and restitution; the mass from its
// inverse and its inertias.
HD void parameters_world(const Trajectory& T, int wld, double* mass_b, double* friction_b, double* restitution_b) {
  const Bodies& B = T.B;
  const int boff = T.body_offset[wld], nb = T.body_offset[wld + 1] - boff;
  for (int s = 0; s < nb; s++) { friction_b[boff + s] = 0; restitution_b[boff + s] = 0; mass_b[boff + s] = 0; }
  for (int t = 0; t < T.steps; t++) {
    const int r0 = T.row_offset[t * T.worlds + wld], r1 = T.row_offset[t * T.worlds + wld + 1];
    for (int r = r0; r < r1; r++) {
      const int ga = boff + T.R.a[r], gb = boff + T.R.b[r];
      const double fa = B.friction[ga], fb = B.friction[gb];
      const double mu = dsqrt(fa * fb);
      if (mu > 0) {
        friction_b[ga] += T.RB.mu[r] * fb / (2 * mu);
        friction_b[gb] += T.RB.mu[r] * fa / (2 * mu);
      }
      if (B.restitution[gb] > B.restitution[ga]) restitution_b[gb] += T.RB.e[r];
      else restitution_b[ga] += T.RB.e[r];
    }
  }
  for (int s = 0; s < nb; s++) {
    const int g = boff + s;
    if (B.type[g] != DYNAMIC) continue;
    const double m = B.mass[g], m0 = B.mass0[g];
    double mb = -T.BB.inv_mass[g] * B.inv_mass[g] / m;
    for (int k = 0; k < 3; k++) {
      mb += T.BB.inertia[3 * g + k] * B.inertia0[3 * g + k] / m0;
      mb -= T.BB.inv_inertia[3 * g + k] * B.inv_inertia[3 * g + k] / m;
    }
    mass_b[g] = mb;
  }
}

}  // namespace adj
}  // namespace bm
