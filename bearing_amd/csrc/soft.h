// Device-side view of a batch (the columns of bearing_mi355x.h, in HBM).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BM_STARTS_DEV 35  // colours, overflow, resting, end (src/contact.js:160)

namespace bm {

struct SoftDev {
  int n_worlds;
  const int *body_offset, *manifold_offset, *pair_offset, *awake_offset;
  const int *starts, *solved, *substeps;
  const double *dt, *gravity;
  double *position, *rotation, *velocity, *angular, *start, *start_q, *inv_i, *tumbled, *ceiling;
  const double *dims, *inv_mass, *inertia, *inv_inertia, *axes, *gravity_scale, *damping, *angular_damping;
  const uint8_t *type, *shape, *frozen, *lock, *framed, *awake;
  const int *joint_head, *awake_list;
  const int *a, *b, *pair, *sub, *row, *colour, *order;
  const uint8_t* points;
  uint8_t* sliding;
  const double *normal, *t1, *t2, *friction, *restitution, *tolerance, *patch;
  double *roll_impulse1, *roll_impulse2, *spin_impulse, *tangent_impulse1, *tangent_impulse2, *drift1, *drift2;
  const double *separation, *point, *shift;
  double *normal_impulse, *max_normal_impulse, *relative_velocity, *gap, *now, *ra, *rb, *la, *lb, *adjusted,
      *normal_mass, *jac, *block, *lever, *anchor_a, *anchor_b, *fjac, *tangent_mass, *roll_mass1, *roll_mass2,
      *spin_mass;
  const uint32_t* pid;
  // What each solved row of a pair's own manifold stores for its pair (src/contact.js:1067-1088
  // store), by row; the download writes it into the pair columns at the row's pair.
  uint8_t* pair_points;
  uint32_t* pair_pid;
  double *pair_normal_impulse, *pair_twist, *pair_friction, *pair_slip;
  // The order workgroups take the worlds in (null: as numbered): the largest first, so the
  // longest worlds of an uneven batch start first and the launch ends with short ones.
  const int* world_order;
};

// One workgroup per world; lds_bytes > 0 keeps each world's body velocities in that
// much dynamic LDS (6 doubles a body) for the phase.
// shape: 0 = 256 threads, registers for one wave a SIMD; 1 = 256 threads, two waves;
// 2 = 512 threads, two waves.
constexpr int SOFT_SHAPES = 3;
void launch_soft_step(const SoftDev& d, int grid, int shape, size_t lds_bytes, hipStream_t stream);
// World wld's soft step across a cooperative grid sized for `work` (its manifolds or bodies,
// whichever is more); the workgroups launched, or -1 if the launch failed. `words`: the grid
// barrier's counters, WIDE_WORDS of them in device memory, zero before the first launch (each
// launch leaves them so); launches sharing them must not overlap.
constexpr int WIDE_GROUPS = 8, WIDE_STRIDE = 32;  // counters 128 bytes apart
constexpr int WIDE_WORDS = WIDE_STRIDE * (WIDE_GROUPS + 2);
hipError_t launch_soft_wide(const SoftDev& d, int wld, int work, unsigned* words, int* pick, int device,
                            hipStream_t stream, int* blocks_out);
size_t soft_step_lds_limit(int device);

}  // namespace bm
