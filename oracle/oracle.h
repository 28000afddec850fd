// ORACLE — test infrastructure only (see jsmath.h). The C interface of the CPU
// restatement: one struct of the reference's own columns, by slot / manifold /
// point / pair, exactly as src/world.js and src/contact.js lay them out.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleSoft {
  int32_t top, awake_count, substeps, count, solved, pair_count;
  double dt, gravity[3];
  /* bodies, by slot: src/world.js:228-311 */
  double *position, *rotation, *velocity, *angular, *dims, *inv_mass, *inertia, *inv_inertia, *inv_i, *axes,
      *tumbled, *ceiling, *gravity_scale, *damping, *angular_damping, *start, *start_q;
  uint8_t *type, *shape, *frozen, *lock, *framed, *awake;
  int32_t *joint_head, *awake_list;
  /* manifolds (m) and their points (4m + k): src/contact.js:149-258 */
  int32_t *starts, *a, *b, *pair, *sub, *row;
  uint8_t *points, *sliding;
  double *normal, *t1, *t2, *friction, *restitution, *tolerance, *patch, *roll_impulse1, *roll_impulse2,
      *spin_impulse, *tangent_impulse1, *tangent_impulse2, *drift1, *drift2, *separation, *point, *shift,
      *normal_impulse, *max_normal_impulse, *relative_velocity, *gap, *now, *ra, *rb, *la, *lb, *adjusted,
      *normal_mass, *jac, *block, *lever, *anchor_a, *anchor_b, *fjac, *tangent_mass, *roll_mass1, *roll_mass2,
      *spin_mass;
  uint32_t *pid;
  /* what the pair list keeps per pair: src/pairs.js:40-110 */
  uint8_t *pair_points;
  uint32_t *pair_pid;
  double *pair_normal_impulse, *pair_twist, *pair_friction, *pair_slip;
} OracleSoft;

/* src/world.js:961-1031: inertias, start poses, prepare, makeSoft, the substeps
   (accelerate, warm start, solve, move, relax), restitution, store.
   Returns 0, or -1 for a manifold this restatement does not cover (a mesh's or
   a compound's, sub >= 0). */
int oracle_soft_step(OracleSoft* w);

/* oracle_soft_step with `workers` more threads (the reference's workers, src/core/jobs.js,
   src/world.js:173-186): prepare, and each colour's warm start and solve, cut across them;
   the rest on the calling thread. The same bits as oracle_soft_step. */
int oracle_soft_step_workers(OracleSoft* w, int workers);

/* src/graph.js:105-188: each manifold's colour (COLOURS = 32, 32 = overflow, 33 =
   resting), the rows' order (row o holds manifold order[o]), the way back (row), the
   35 run starts, and solved (the resting run's start). count manifolds between slots
   A[m] and B[m]; top slots in use. */
int oracle_colour(int32_t count, const int32_t* A, const int32_t* B, int32_t top, const uint8_t* type,
                  const uint8_t* frozen, int32_t frozen_count, int32_t* colour, int32_t* order, int32_t* row,
                  int32_t* starts, int32_t* solved);

#ifdef __cplusplus
}
#endif
