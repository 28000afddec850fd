// The broadphase's per-body and per-pair work, written once for the device kernels
// (broad.hip) and the host: the bounds of an awake body (src/world.js:1574-1630 _bounds,
// _boundsOf), whether they still lie in its fat box (src/broadphase.js:140-159 update's
// test), a child1-first depth-first walk of an AABB tree for the leaves overlapping a box
// (src/tree.js:259-275 query) and the fat-box overlap of a pair (src/broadphase.js:181-187,
// src/pairs.js:340-356 holds). Every expression keeps the reference's order of operations:
// sources that include this are built with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jsmath.h"

namespace bm {
namespace broad {

constexpr double PIXEL = 1.0 / 200;  // src/contact.js:125
constexpr int DYNAMIC = 0, STATIC = 1;
constexpr int BOX = 0, SPHERE = 1, CAPSULE = 2, HULL = 3, MESH = 5;
constexpr int STATIC_TREE = 0, MOVING_TREE = 1;
// The deepest a walk's stack may grow. The trees are AVL-balanced (src/tree.js:219-252), so a
// tree of 2^24 leaves stays within about 1.44 * 24 + 2 levels, and a walk holds at most one
// entry a level and one more; a deeper tree is reported.
constexpr int STACK = 64;

// What the bounds read of each body, by slot. `local`: for a hull or a heightfield, the centre
// (3) and half extents (3) of the box round its corners or samples in its own frame.
struct Bodies {
  const uint8_t *type, *shape;
  const double *dims, *gravity_scale, *reach, *size, *local;
  const double *position, *rotation, *velocity, *angular;
};

// A node of an AABB tree as src/tree.js keeps it, in one 64-byte record: its box (min x y z,
// max x y z), children, height (0 for a leaf) and a leaf's item (its body's slot).
struct Node {
  double box[6];
  int32_t child1, child2, height, item;
};
static_assert(sizeof(Node) == 64, "a node is one 64-byte record");

struct Tree {
  int root;
  const Node* node;
};

// src/core/math.js:223-228 (axesOf)
__host__ __device__ __forceinline__ void axes_of(const double* q, double* ax) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  ax[0] = 1 - 2 * (y * y + z * z); ax[1] = 2 * (x * y + z * w); ax[2] = 2 * (x * z - y * w);
  ax[3] = 2 * (x * y - z * w); ax[4] = 1 - 2 * (x * x + z * z); ax[5] = 2 * (y * z + x * w);
  ax[6] = 2 * (x * z + y * w); ax[7] = 2 * (y * z - x * w); ax[8] = 1 - 2 * (x * x + y * y);
}

// src/world.js:1580-1630 (_boundsOf) for body s, stepping dt under gravity of length gl:
// out[0..8] its axes, out[9] its motion, out[10..15] its box grown by motion and a pixel.
__host__ __device__ __forceinline__ void bounds(const Bodies& b, int s, double dt, double gl, double* out) {
  const int s3 = 3 * s;
  double* ax = out;
  axes_of(b.rotation + 4 * s, ax);
  double m = 0;
  if (b.type[s] != STATIC) {
    const double* v = b.velocity + s3;
    const double* om = b.angular + s3;
    const double speed = dsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        + (b.type[s] == DYNAMIC ? gl * dabs(b.gravity_scale[s]) * dt : 0)
        + dsqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]) * b.reach[s];
    m = speed * dt;
  }
  out[9] = m;
  const int shape = b.shape[s];
  double ex, ey, ez, ox = 0, oy = 0, oz = 0;
  if (shape == BOX || shape == HULL || shape >= MESH) {
    double hx = b.dims[s3] * 0.5, hy = b.dims[s3 + 1] * 0.5, hz = b.dims[s3 + 2] * 0.5;
    if (shape != BOX) {
      // The box round its corners (or samples) in its own frame, turned into the world's.
      const double* l = b.local + 6 * s;
      const double cx = l[0], cy = l[1], cz = l[2];
      hx = l[3]; hy = l[4]; hz = l[5];
      ox = ax[0] * cx + ax[3] * cy + ax[6] * cz;
      oy = ax[1] * cx + ax[4] * cy + ax[7] * cz;
      oz = ax[2] * cx + ax[5] * cy + ax[8] * cz;
    }
    ex = dabs(ax[0]) * hx + dabs(ax[3]) * hy + dabs(ax[6]) * hz;
    ey = dabs(ax[1]) * hx + dabs(ax[4]) * hy + dabs(ax[7]) * hz;
    ez = dabs(ax[2]) * hx + dabs(ax[5]) * hy + dabs(ax[8]) * hz;
  } else if (shape == SPHERE) {
    ex = ey = ez = b.dims[s3];
  } else {
    const double r = b.dims[s3], hl = (b.dims[s3 + 1] - 2 * r) * 0.5;
    ex = dabs(ax[3]) * hl + r; ey = dabs(ax[4]) * hl + r; ez = dabs(ax[5]) * hl + r;
  }
  const double grow = m + b.size[s] * PIXEL;
  const double* p = b.position + s3;
  const double cx = p[0] + ox, cy = p[1] + oy, cz = p[2] + oz;
  double* box = out + 10;
  box[0] = cx - ex - grow; box[1] = cy - ey - grow; box[2] = cz - ez - grow;
  box[3] = cx + ex + grow; box[4] = cy + ey + grow; box[5] = cz + ez + grow;
}

// src/broadphase.js:146-150: box b still lies within fat box f.
__host__ __device__ __forceinline__ bool contains(const double* f, const double* b) {
  return b[0] >= f[0] && b[1] >= f[1] && b[2] >= f[2] && b[3] <= f[3] && b[4] <= f[4] && b[5] <= f[5];
}

// src/broadphase.js:181-187 (overlap) of two fat boxes.
__host__ __device__ __forceinline__ bool overlap(const double* fa, const double* fb) {
  return fa[0] <= fb[3] && fa[3] >= fb[0] && fa[1] <= fb[4] && fa[4] >= fb[1] && fa[2] <= fb[5] && fa[5] >= fb[2];
}

// src/tree.js:259-275 (query): visit(item) for each leaf whose box overlaps q, child1's side
// first. The walk's stack holds STACK entries, entry k at stack[k * stride]; false if the walk
// needed more (it then stops).
template <class F>
__host__ __device__ __forceinline__ bool query(const Tree& t, const double* q, int* stack, int stride, F&& visit) {
  if (t.root < 0) return true;
  stack[0] = t.root;
  int top = 1;
  while (top > 0) {
    const Node& n = t.node[stack[--top * stride]];
    const double* b = n.box;
    if (b[0] > q[3] || b[3] < q[0] || b[1] > q[4] || b[4] < q[1] || b[2] > q[5] || b[5] < q[2]) continue;
    if (n.height == 0) {
      visit(n.item);
      continue;
    }
    if (top + 2 > STACK) return false;
    stack[top++ * stride] = n.child2;
    stack[top++ * stride] = n.child1;
  }
  return true;
}

}  // namespace broad
}  // namespace bm
