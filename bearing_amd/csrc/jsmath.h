// JavaScript's double arithmetic on the device, as the reference relies on it:
// + - * / and sqrt correctly rounded and never fused (every source here is built
// with -ffp-contract=off), Math.max with its signed-zero rule, and fdlibm's
// atan / atan2 (Sun Microsystems, 1993; the reference restates them in
// src/core/math.js:120-200) with the same breakpoints and coefficients, so the
// device gets the reference's bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {

// Math.max: NaN if either is NaN, +0 over -0 (src/contact.js:1866 relies on it).
__host__ __device__ __forceinline__ double js_max(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a > b) return a;
  if (b > a) return b;
  return __builtin_signbit(a) ? b : a;
}

__host__ __device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__host__ __device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }

// fdlibm s_atan.c: reduce about 0, 0.5, 1, 1.5 or infinity, then an odd polynomial.
__host__ __device__ __forceinline__ double js_atan(double x) {
  // The arctangents of 0.5, 1, 1.5 and infinity, each in a high and a low part; picked
  // by branches rather than a table, which the compiler would put in scratch memory.
  const double HI3 = 1.57079632679489655800e+00, LO3 = 6.12323399573676603587e-17;
  if (x != x) return x;
  const double ax = dabs(x);
  if (ax >= 7.378697629483821e+19) return x > 0 ? HI3 + LO3 : -HI3 - LO3;
  int id;
  double t;
  if (ax < 0.4375) {
    if (ax < 7.450580596923828e-09) return x;
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
  const double s1 = z * (3.33333333333329318027e-01 + w * (1.42857142725034663711e-01
      + w * (9.09088713343650656196e-02 + w * (6.66107313738753120669e-02
      + w * (4.97687799461593236017e-02 + w * 1.62858201153657823623e-02)))));
  const double s2 = w * (-1.99999999998764832476e-01 + w * (-1.11111104054623557880e-01
      + w * (-7.69187620504482999495e-02 + w * (-5.83357013379057348645e-02
      + w * -3.65315727442169155270e-02))));
  if (id < 0) return t - t * (s1 + s2);
  const double hi = id == 0 ? 4.63647609000806093515e-01 : id == 1 ? 7.85398163397448278999e-01
      : id == 2 ? 9.82793723247329054082e-01 : HI3;
  const double lo = id == 0 ? 2.26987774529616870924e-17 : id == 1 ? 3.06161699786838301793e-17
      : id == 2 ? 1.39033110312309984516e-17 : LO3;
  const double r = hi - ((t * (s1 + s2) - lo) - t);
  return x < 0 ? -r : r;
}

// fdlibm e_atan2.c: the quadrant from the signs (-0 included), then atan of |y/x|.
__host__ __device__ __forceinline__ double js_atan2(double y, double x) {
  const double PI = 3.1415926535897931160e+00, PI_LO = 1.2246467991473531772e-16;
  const double PI_O_2 = 1.5707963267948965580e+00, PI_O_4 = 7.8539816339744827900e-01;
  if (x != x || y != y) return x + y;
  if (x == 1) return js_atan(y);
  int m = (__builtin_signbit(y) ? 1 : 0) | (__builtin_signbit(x) ? 2 : 0);
  if (y == 0) return m < 2 ? y : m == 2 ? PI : -PI;
  if (x == 0) return y < 0 ? -PI_O_2 : PI_O_2;
  const double ax = dabs(x), ay = dabs(y);
  if (ax == __builtin_inf()) {
    if (ay == __builtin_inf()) return m == 0 ? PI_O_4 : m == 1 ? -PI_O_4 : m == 2 ? 3 * PI_O_4 : -3 * PI_O_4;
    return m == 0 ? 0.0 : m == 1 ? -0.0 : m == 2 ? PI : -PI;
  }
  if (ay == __builtin_inf()) return y < 0 ? -PI_O_2 : PI_O_2;
  double z;
  if (ay > ax * 1.152921504606847e+18) {
    z = PI_O_2 + 0.5 * PI_LO;
    m &= 1;
  } else if (x < 0 && ay < ax * 8.673617379884035e-19) {
    z = 0;
  } else {
    z = js_atan(ay / ax);
  }
  switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return PI - (z - PI_LO);
    default: return (z - PI_LO) - PI;
  }
}

// Sine and cosine as src/core/math.js:26-123 computes them, fdlibm's published method
// (k_sin.c, k_cos.c, the medium case of e_rem_pio2.c): x less the nearest multiple n of
// pi/2 in 33-bit pieces (each product with n exact), as a head and a tail, then the
// minimax polynomials on [-pi/4, pi/4], chosen and signed by n mod 4. Math.round's
// halves-up rounding of x * 2/pi is floor, then up when the remainder is at least 0.5.
struct Reduced {
  double head, tail;
  int quadrant;
};
__host__ __device__ __forceinline__ Reduced reduce_pio2(double x) {
  const double n0 = __builtin_floor(x * 6.36619772367581382433e-01);
  const double n = x * 6.36619772367581382433e-01 - n0 >= 0.5 ? n0 + 1 : n0;
  double r = x - n * 1.57079632673412561417e+00;
  double w = n * 6.07710050650619224932e-11;
  double head = r - w;
  const double limit = dabs(x);
  if (dabs(head) < limit * 3.0517578125e-05) {
    double t = r;
    w = n * 6.07710050630396597660e-11;
    r = t - w;
    w = n * 2.02226624879595063154e-21 - ((t - r) - w);
    head = r - w;
    if (dabs(head) < limit * 8.881784197001252e-16) {
      t = r;
      w = n * 2.02226624871116645580e-21;
      r = t - w;
      w = n * 8.47842766036889956997e-32 - ((t - r) - w);
      head = r - w;
    }
  }
  Reduced out;
  out.head = head;
  out.tail = (r - head) - w;
  out.quadrant = (int)((long long)n & 3);
  return out;
}
__host__ __device__ __forceinline__ double kernel_sin(double x, double y) {
  const double z = x * x, v = z * x;
  const double r = 8.33333333332248946124e-03 +
      z * (-1.98412698298579493134e-04 + z * (2.75573137070700676789e-06 +
      z * (-2.50507602534068634195e-08 + z * 1.58969099521155010221e-10)));
  return x - ((z * (0.5 * y - v * r) - y) - v * -1.66666666666666324348e-01);
}
__host__ __device__ __forceinline__ double kernel_cos(double x, double y) {
  const double z = x * x;
  const double r = z * (4.16666666666666019037e-02 + z * (-1.38888888888741095749e-03 +
      z * (2.48015872894767294178e-05 + z * (-2.75573143513906633035e-07 +
      z * (2.08757232129817482790e-09 + z * -1.13596475577881948265e-11)))));
  const double hz = 0.5 * z, w = 1 - hz;
  return w + (((1 - w) - hz) + (z * r - x * y));
}
__host__ __device__ __forceinline__ double js_sin(double x) {
  const Reduced r = reduce_pio2(x);
  switch (r.quadrant) {
    case 0: return kernel_sin(r.head, r.tail);
    case 1: return kernel_cos(r.head, r.tail);
    case 2: return -kernel_sin(r.head, r.tail);
    default: return -kernel_cos(r.head, r.tail);
  }
}
__host__ __device__ __forceinline__ double js_cos(double x) {
  const Reduced r = reduce_pio2(x);
  switch (r.quadrant) {
    case 0: return kernel_cos(r.head, r.tail);
    case 1: return -kernel_sin(r.head, r.tail);
    case 2: return -kernel_cos(r.head, r.tail);
    default: return kernel_sin(r.head, r.tail);
  }
}

// Rotation by a unit quaternion (x, y, z, w), or by its conjugate:
// v' = v + w t + u x t, t = 2 u x v (src/core/math.js:212-220).
struct V3 {
  double x, y, z;
};

__host__ __device__ __forceinline__ V3 rotate(double qx, double qy, double qz, double qw, double x, double y, double z,
                                     bool inverse) {
  const double s = inverse ? -1 : 1;
  const double ux = qx * s, uy = qy * s, uz = qz * s, w = qw;
  const double tx = 2 * (uy * z - uz * y), ty = 2 * (uz * x - ux * z), tz = 2 * (ux * y - uy * x);
  V3 o;
  o.x = x + w * tx + (uy * tz - uz * ty);
  o.y = y + w * ty + (uz * tx - ux * tz);
  o.z = z + w * tz + (ux * ty - uy * tx);
  return o;
}

__host__ __device__ __forceinline__ V3 rotate(const double* q, double x, double y, double z, bool inverse) {
  return rotate(q[0], q[1], q[2], q[3], x, y, z, inverse);
}

}  // namespace bm
