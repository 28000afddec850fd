// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg may load it, and only as the checker.
//
// JavaScript's arithmetic as the reference uses it, in plain C++ for the CPU
// restatement: IEEE doubles, + - * / and sqrt correctly rounded, never fused
// (this folder is compiled with -ffp-contract=off), and Math.min / Math.max with
// their NaN and signed-zero rules. atan and atan2 follow fdlibm's published
// s_atan.c / e_atan2.c (Sun Microsystems, 1993), the method the reference
// restates in src/core/math.js:120-200.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

// Math.max(a, b): NaN if either is NaN; +0 over -0.
inline double js_max(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

// Math.min(a, b): NaN if either is NaN; -0 over +0.
inline double js_min(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}

// src/core/math.js:124-133, fdlibm's atan breakpoints and coefficients.
static const double ATAN_HI[4] = {4.63647609000806093515e-01, 7.85398163397448278999e-01,
                                  9.82793723247329054082e-01, 1.57079632679489655800e+00};
static const double ATAN_LO[4] = {2.26987774529616870924e-17, 3.06161699786838301793e-17,
                                  1.39033110312309984516e-17, 6.12323399573676603587e-17};
static const double AT[11] = {3.33333333333329318027e-01, -1.99999999998764832476e-01,
                              1.42857142725034663711e-01, -1.11111104054623557880e-01,
                              9.09088713343650656196e-02, -7.69187620504482999495e-02,
                              6.66107313738753120669e-02, -5.83357013379057348645e-02,
                              4.97687799461593236017e-02, -3.65315727442169155270e-02,
                              1.62858201153657823623e-02};
static const double PI_ = 3.1415926535897931160e+00, PI_LO = 1.2246467991473531772e-16;
static const double PI_O_2 = 1.5707963267948965580e+00, PI_O_4 = 7.8539816339744827900e-01;
static const double TWO_66 = 7.378697629483821e+19, TWO_M27 = 7.450580596923828e-09;
static const double TWO_60 = 1.152921504606847e+18, TWO_M60 = 8.673617379884035e-19;

// src/core/math.js:135-160 (fdlibm s_atan.c): reduce about 0, 0.5, 1, 1.5 or infinity, then an odd polynomial.
inline double js_atan(double x) {
  if (x != x) return x;
  const double ax = std::fabs(x);
  if (ax >= TWO_66) return x > 0 ? ATAN_HI[3] + ATAN_LO[3] : -ATAN_HI[3] - ATAN_LO[3];
  int id;
  double t;
  if (ax < 0.4375) {
    if (ax < TWO_M27) return x;
    id = -1; t = x;
  } else if (ax < 0.6875) {
    id = 0; t = (2 * ax - 1) / (2 + ax);
  } else if (ax < 1.1875) {
    id = 1; t = (ax - 1) / (ax + 1);
  } else if (ax < 2.4375) {
    id = 2; t = (ax - 1.5) / (1 + 1.5 * ax);
  } else {
    id = 3; t = -1 / ax;
  }
  const double z = t * t, w = z * z;
  const double s1 = z * (AT[0] + w * (AT[2] + w * (AT[4] + w * (AT[6] + w * (AT[8] + w * AT[10])))));
  const double s2 = w * (AT[1] + w * (AT[3] + w * (AT[5] + w * (AT[7] + w * AT[9]))));
  if (id < 0) return t - t * (s1 + s2);
  const double r = ATAN_HI[id] - ((t * (s1 + s2) - ATAN_LO[id]) - t);
  return x < 0 ? -r : r;
}

// src/core/math.js:172-200 (fdlibm e_atan2.c), the quadrant from the two signs, -0 included.
inline double js_atan2(double y, double x) {
  if (x != x || y != y) return x + y;
  if (x == 1) return js_atan(y);
  int m = (std::signbit(y) ? 1 : 0) | (std::signbit(x) ? 2 : 0);
  if (y == 0) return m < 2 ? y : m == 2 ? PI_ : -PI_;
  if (x == 0) return y < 0 ? -PI_O_2 : PI_O_2;
  const double ax = std::fabs(x), ay = std::fabs(y);
  if (ax == INFINITY) {
    if (ay == INFINITY) {
      const double r[4] = {PI_O_4, -PI_O_4, 3 * PI_O_4, -3 * PI_O_4};
      return r[m];
    }
    const double r[4] = {0.0, -0.0, PI_, -PI_};
    return r[m];
  }
  if (ay == INFINITY) return y < 0 ? -PI_O_2 : PI_O_2;
  double z;
  if (ay > ax * TWO_60) {
    z = PI_O_2 + 0.5 * PI_LO;
    m &= 1;
  } else if (x < 0 && ay < ax * TWO_M60) {
    z = 0;
  } else {
    z = js_atan(ay / ax);
  }
  switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return PI_ - (z - PI_LO);
    default: return (z - PI_LO) - PI_;
  }
}

// Math.round(x): the nearest integer, halves toward +infinity, -0 for -0.5 <= x < 0.
inline double js_round(double x) {
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1;
  if (r == 0 && (x < 0 || std::signbit(x))) return -0.0;
  return r;
}

// Sine and cosine as src/core/math.js:26-123 computes them, fdlibm's published method
// (k_sin.c, k_cos.c and the medium case of e_rem_pio2.c, Sun Microsystems, 1993):
// x less the nearest multiple n of pi/2, taken off in 33-bit pieces of pi/2 so each
// product with n is exact, as a head and a tail; then the minimax polynomials for
// sine or cosine on [-pi/4, pi/4], chosen and signed by n mod 4.
struct Reduced {
  double head, tail;
  int quadrant;
};
inline Reduced reduce_pio2(double x) {
  const double INV_PIO2 = 6.36619772367581382433e-01;
  const double PIO2_1 = 1.57079632673412561417e+00, PIO2_1T = 6.07710050650619224932e-11;
  const double PIO2_2 = 6.07710050630396597660e-11, PIO2_2T = 2.02226624879595063154e-21;
  const double PIO2_3 = 2.02226624871116645580e-21, PIO2_3T = 8.47842766036889956997e-32;
  const double ROUND_2 = 3.0517578125e-05, ROUND_3 = 8.881784197001252e-16;  // 2^-15, 2^-50
  const double n = js_round(x * INV_PIO2);
  double r = x - n * PIO2_1;
  double w = n * PIO2_1T;
  double head = r - w;
  const double limit = std::fabs(x);
  if (std::fabs(head) < limit * ROUND_2) {
    double t = r;
    w = n * PIO2_2;
    r = t - w;
    w = n * PIO2_2T - ((t - r) - w);
    head = r - w;
    if (std::fabs(head) < limit * ROUND_3) {
      t = r;
      w = n * PIO2_3;
      r = t - w;
      w = n * PIO2_3T - ((t - r) - w);
      head = r - w;
    }
  }
  Reduced out;
  out.head = head;
  out.tail = (r - head) - w;
  // n & 3 of a JavaScript number: ToInt32, two's complement.
  out.quadrant = (int)((int64_t)n & 3);
  return out;
}
inline double kernel_sin(double x, double y) {
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03;
  const double S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06;
  const double S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  const double z = x * x, v = z * x;
  const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}
inline double kernel_cos(double x, double y) {
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03;
  const double C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07;
  const double C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  const double z = x * x;
  const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  const double hz = 0.5 * z, w = 1 - hz;
  return w + (((1 - w) - hz) + (z * r - x * y));
}
inline double js_sin(double x) {
  const Reduced r = reduce_pio2(x);
  switch (r.quadrant) {
    case 0: return kernel_sin(r.head, r.tail);
    case 1: return kernel_cos(r.head, r.tail);
    case 2: return -kernel_sin(r.head, r.tail);
    default: return -kernel_cos(r.head, r.tail);
  }
}
inline double js_cos(double x) {
  const Reduced r = reduce_pio2(x);
  switch (r.quadrant) {
    case 0: return kernel_cos(r.head, r.tail);
    case 1: return -kernel_sin(r.head, r.tail);
    case 2: return -kernel_cos(r.head, r.tail);
    default: return kernel_sin(r.head, r.tail);
  }
}

// src/core/math.js:212-220: v' = v + w t + u x t with t = 2 u x v; `inverse` uses the conjugate.
inline void rotate(const double* q, double x, double y, double z, bool inverse, double* out) {
  const double s = inverse ? -1 : 1;
  const double ux = q[0] * s, uy = q[1] * s, uz = q[2] * s, w = q[3];
  const double tx = 2 * (uy * z - uz * y), ty = 2 * (uz * x - ux * z), tz = 2 * (ux * y - uy * x);
  out[0] = x + w * tx + (uy * tz - uz * ty);
  out[1] = y + w * ty + (uz * tx - ux * tz);
  out[2] = z + w * tz + (ux * ty - uy * tx);
}

// src/core/math.js:223-228: the rotation's matrix columns.
inline void axes_of(const double* q, double* ax) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  ax[0] = 1 - 2 * (y * y + z * z); ax[1] = 2 * (x * y + z * w); ax[2] = 2 * (x * z - y * w);
  ax[3] = 2 * (x * y - z * w); ax[4] = 1 - 2 * (x * x + z * z); ax[5] = 2 * (y * z + x * w);
  ax[6] = 2 * (x * z + y * w); ax[7] = 2 * (y * z - x * w); ax[8] = 1 - 2 * (x * x + y * y);
}

// src/core/math.js:231-238: out = a (x) b.
inline void quat_multiply(const double* a, const double* b, double* out) {
  const double ax = a[0], ay = a[1], az = a[2], aw = a[3];
  const double bx = b[0], by = b[1], bz = b[2], bw = b[3];
  out[0] = aw * bx + ax * bw + ay * bz - az * by;
  out[1] = aw * by + ay * bw + az * bx - ax * bz;
  out[2] = aw * bz + az * bw + ax * by - ay * bx;
  out[3] = aw * bw - ax * bx - ay * by - az * bz;
}

}  // namespace oracle
