// Device-side view of the broadphase passes (broad.hip): the bounds and fat-box test of awake
// bodies, tree walks for the bodies the update marked, and the fat-box overlap of pairs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "broad_dev.h"

namespace bm {

// n awake bodies list[0..n): out field-major, out[k * n + i] for k in 0..15 (axes 9, motion,
// box 6, as broad::bounds), and refile[i] = 1 where the body is in the moving tree and its new
// box left its leaf's fat box.
struct BoundsDev {
  int n;
  const int* list;
  double dt, gl;
  broad::Bodies bodies;
  broad::Tree moving;
  const int* leaf;
  const int8_t* which;
  double* out;
  uint8_t* refile;
};
void launch_bounds(const BoundsDev& d, hipStream_t stream);

// n marked bodies moved[0..n): each one's leaf box walked down the static tree (unless the body
// is static) and then the moving tree; counts[i] the leaves met other than itself, starts[i]
// where they were written in hits (room for cap), in walk order. total: the hits of every body
// (reserved in any order); flags |= 1 when a walk outgrew its stack.
struct QueryDev {
  int n;
  const int* moved;
  broad::Tree trees[2];
  const int* leaf;
  const int8_t* which;
  const uint8_t* type;
  int *counts, *starts, *hits;
  int cap;
  int *total, *flags;
};
void launch_query(const QueryDev& d, hipStream_t stream);

// n pairs (a[i], b[i]) by slot: apart[i] = 1 where their leaves' fat boxes no longer overlap.
struct HoldsDev {
  int n;
  const int *a, *b;
  broad::Tree trees[2];
  const int* leaf;
  const int8_t* which;
  uint8_t* apart;
};
void launch_holds(const HoldsDev& d, hipStream_t stream);

// rows[13 i ..] (position 3, rotation 4, velocity 3, angular 3) into slot slots[i] of the four
// per-slot arrays, for n rows.
void launch_broad_scatter(int n, const int* slots, const double* rows, double* position, double* rotation,
                          double* velocity, double* angular, hipStream_t stream);

}  // namespace bm
