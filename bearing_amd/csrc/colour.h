// Device-side views for colouring, arranging rows and rest() (colour.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "soft.h"

namespace bm {

struct ColourDev {
  int n_worlds;
  const int *body_offset, *manifold_offset;
  const int *a, *b;                // narrow order, each world's own slots
  const uint8_t *type, *frozen;
  int* colour;                     // out, narrow order
  int *order, *row, *starts, *solved;
  int* deg;                        // scratch: bodies + n_worlds
  int* adj;                        // scratch: 2 per manifold
  uint64_t* keys;                  // scratch: key_offset[n_worlds]
  const int* key_offset;           // per world, a power of two at least its manifolds
  int in_lds;                      // 1: each world's colouring and sort run in dynamic LDS
  const int* world_order;          // the order workgroups take the worlds in (null: as numbered)
};

// Dynamic LDS a world takes when coloured in LDS: first the held masks, the manifolds' ends
// and the moving flags, which the sort keys then overwrite, then the colours.
__host__ __device__ inline size_t colour_lds_front(int bodies, int manifolds, int width) {
  const size_t colouring = 4 * (size_t)(bodies + (bodies & 1)) + 8 * (size_t)manifolds + (size_t)bodies;
  const size_t sorting = 8 * (size_t)width;
  return ((colouring > sorting ? colouring : sorting) + 15) & ~(size_t)15;
}
inline size_t colour_lds_bytes(int bodies, int manifolds, int width) {
  return colour_lds_front(bodies, manifolds, width) + 4 * (size_t)manifolds;
}
size_t colour_lds_limit(int device);

struct ArrangeColumn {
  const void* from;
  void* to;
  int bytes;  // 1, 4 or 8
  int per;    // numbers per manifold
};

constexpr int ARRANGE_MAX = 32;

struct ArrangeDev {
  long manifolds;
  const int *manifold_offset, *order, *world_of;
  ArrangeColumn columns[ARRANGE_MAX];
};

struct RestDev {
  long manifolds;
  const int *manifold_offset, *solved, *world_of;
  const uint8_t* points;
  const double *separation, *normal_impulse;
  double *max_normal_impulse, *relative_velocity, *gap;
};

void launch_colour(const ColourDev& d, size_t lds_bytes, hipStream_t stream);
// One world (the batch's wld-th, at manifold offset 0) coloured by streaming its manifolds
// through one wave with the bodies' held masks in LDS, its rows ordered by a radix sort:
// for a world too large to colour in LDS. LDS it takes; sort scratch it needs; 0, or -1 if
// the sort could not run.
size_t colour_stream_lds(int bodies);
size_t colour_stream_scratch(int manifolds);
int launch_colour_stream(const ColourDev& d, int wld, int manifolds, int bodies, void* scratch, size_t scratch_bytes,
                         hipStream_t stream);
void launch_world_of(const int* manifold_offset, int n_worlds, int* world_of, hipStream_t stream);
void launch_arrange(const ArrangeDev& d, int columns, hipStream_t stream);
void launch_rest(const RestDev& d, hipStream_t stream);

}  // namespace bm
