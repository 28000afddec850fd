// Device-side view of a batch of shape pairs for the narrowphase (collide.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {

struct CollideDev {
  int n;
  const double *a, *b;             // shape records: kind, centre 3, axes 9, half 3, weight 3
  const double *margin, *tolerance;
  int* count;                      // -1 for a pair of kinds not handled (capsules, hulls)
  double *normal, *points, *separations;
  uint32_t* ids;
};

void launch_collide(const CollideDev& d, hipStream_t stream);

struct MeetDev {
  int n;
  const double *a, *b;  // body records: the shape record (19), rotation 4, velocity 3, spin 3,
                        // gravity scale, reach, flags (1 dynamic, 2 spin capped)
  const double* pair;   // margin, tolerance, dt, gravity 3
  int *count, *flags;   // flags: 1 shifted, 2 a time of impact was asked
  double *normal, *points, *separations, *shift;
  uint32_t* ids;
};

void launch_meet(const MeetDev& d, hipStream_t stream);

}  // namespace bm
