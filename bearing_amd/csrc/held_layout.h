// Where a held trajectory's columns lie in one buffer (the device's, or in the tests the
// host's), and which of them come from the bm_held description. Host code.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <cstring>
#include <vector>

#include "../../include/bearing_mi355x.h"
#include "adjoint.h"

namespace bm {
namespace adj {

struct Copy {
  void* to;
  const void* from;
  size_t bytes;
};

// The per-body parameters and state the forward and backward passes take and give.
struct HeldIO {
  double *state0, *mass_in, *friction_in, *restitution_in, *state_out;   // state: BM_HELD_STATE a body
  double *state_bar, *state0_bar, *mass_bar, *friction_bar, *restitution_bar;
  double *left0, *ceiling0;
};

// Where each (step, world)'s rows start among the working rows: each world has a block of
// its own, as long as its longest segment of `every` steps, so worlds at different steps
// (their workgroups run apart) never share a working row. Returns the working rows in all.
inline int work_blocks(const bm_held& h, std::vector<int32_t>& at) {
  at.assign((size_t)h.steps * h.worlds, 0);
  int offset = 0;
  for (int w = 0; w < h.worlds; w++) {
    int most = 0, run = 0;
    for (int t = 0; t < h.steps; t++) {
      if (t % h.every == 0) run = 0;
      at[(size_t)t * h.worlds + w] = offset + run;
      run += h.row_offset[t * h.worlds + w + 1] - h.row_offset[t * h.worlds + w];
      if (run > most) most = run;
    }
    offset += most;
  }
  return offset;
}

// Lays the trajectory's columns out from `base` (nullptr: only measure), returning the
// bytes taken; fills T and io with the columns' addresses and `copies` with what comes
// from h.
inline size_t lay_out(const bm_held& h, char* base, Trajectory& T, HeldIO& io, std::vector<Copy>* copies,
                      std::vector<int32_t>& work_at) {
  size_t at = 0;
  auto take = [&](size_t bytes) -> char* {
    at = (at + 255) & ~(size_t)255;
    char* p = base ? base + at : nullptr;
    at += bytes;
    return p;
  };
  auto from = [&](const void* src, size_t bytes) -> char* {
    char* p = take(bytes);
    if (copies && base && bytes) copies->push_back(Copy{p, src, bytes});
    return p;
  };
  const size_t nb = h.bodies, nw = h.worlds, rows = h.rows;
  int most_sub = 1;
  for (int k = 0; k < h.worlds; k++)
    if (h.substeps[k] > most_sub) most_sub = h.substeps[k];
  const int blocks = work_blocks(h, work_at);
  const size_t work = blocks > 0 ? blocks : 1;
  T.worlds = h.worlds; T.steps = h.steps; T.every = h.every; T.bodies = h.bodies;
  T.substeps_most = most_sub; T.work_rows = (int)work;
  auto i32 = [&](const int32_t* src, size_t n) { return reinterpret_cast<int*>(from(src, n * 4)); };
  auto u8 = [&](const uint8_t* src, size_t n) { return reinterpret_cast<uint8_t*>(from(src, n)); };
  auto f64 = [&](const double* src, size_t n) { return reinterpret_cast<double*>(from(src, n * 8)); };
  auto d = [&](size_t n) { return reinterpret_cast<double*>(take(n * 8)); };
  T.body_offset = i32(h.body_offset, nw + 1);
  T.row_offset = i32(h.row_offset, (size_t)h.steps * nw + 1);
  T.starts = i32(h.starts, (size_t)h.steps * nw * STARTS);
  T.narrow = i32(h.narrow, rows);
  T.work_at = i32(work_at.data(), (size_t)h.steps * nw);
  T.substeps = i32(h.substeps, nw);
  T.dt = f64(h.dt, nw);
  T.gravity = f64(h.gravity, 3 * nw);
  Bodies& B = T.B;
  B.type = u8(h.type, nb); B.shape = u8(h.shape, nb); B.tumbles = u8(h.tumbles, nb); B.capped = u8(h.capped, nb);
  B.rounded = f64(h.rounded, nb);
  T.Z.segment = f64(h.segment, nb);
  B.gravity_scale = f64(h.gravity_scale, nb); B.damping = f64(h.damping, nb);
  B.angular_damping = f64(h.angular_damping, nb);
  B.mass0 = f64(h.mass0, nb); B.inv_mass0 = f64(h.inv_mass0, nb);
  B.inertia0 = f64(h.inertia0, 3 * nb); B.inv_inertia0 = f64(h.inv_inertia0, 3 * nb);
  io.left0 = f64(h.left, 3 * nb); io.ceiling0 = f64(h.ceiling, nb);
  io.mass_in = d(nb); io.friction_in = d(nb); io.restitution_in = d(nb);
  B.mass = io.mass_in; B.friction = io.friction_in; B.restitution = io.restitution_in;
  B.inv_mass = d(nb); B.inertia = d(3 * nb); B.inv_inertia = d(3 * nb);
  B.p = d(3 * nb); B.q = d(4 * nb); B.v = d(3 * nb); B.w = d(3 * nb); B.left = d(3 * nb); B.ceiling = d(nb);
  B.axes = nullptr; B.inv_i = d(6 * nb); B.start = d(3 * nb); B.start_q = d(4 * nb);
  io.state0 = d(BM_HELD_STATE * nb); io.state_out = d(BM_HELD_STATE * nb);
  io.state_bar = d(BM_HELD_STATE * nb); io.state0_bar = d(BM_HELD_STATE * nb);
  io.mass_bar = d(nb); io.friction_bar = d(nb); io.restitution_bar = d(nb);
  BodiesBar& BB = T.BB;
  BB.p = d(3 * nb); BB.q = d(4 * nb); BB.v = d(3 * nb); BB.w = d(3 * nb); BB.left = d(3 * nb); BB.ceiling = d(nb);
  BB.inv_i = d(6 * nb); BB.start = d(3 * nb); BB.start_q = d(4 * nb); BB.inv_mass = d(nb);
  BB.inertia = d(3 * nb); BB.inv_inertia = d(3 * nb);
  Rows& R = T.R;
  R.a = i32(h.a, rows); R.b = i32(h.b, rows); R.src = i32(h.src, rows);
  R.count = u8(h.count, rows);
  R.pid = reinterpret_cast<const uint32_t*>(from(h.pid, (size_t)P * rows * 4));
  R.local_normal = f64(h.local_normal, 3 * rows); R.ends = f64(h.ends, 6 * P * rows);
  R.tolerance = f64(h.tolerance, rows); R.patch = f64(h.patch, rows);
  R.shift = f64(h.shift, 3 * P * rows);
  // the stores: the first rows' from h, the rest written by the forward pass
  auto stores = [&](const double* src, size_t per) {
    double* p = d(per * rows);
    if (copies && base && h.first) copies->push_back(Copy{p, src, per * h.first * 8});
    return p;
  };
  R.store_impulse = stores(h.store_impulse, P);
  R.store_friction = stores(h.store_friction, 3);
  R.store_twist = stores(h.store_twist, 3);
  R.store_slip = stores(h.store_slip, 3);
  RowsBar& RB = T.RB;
  RB.store_impulse = d(P * rows); RB.store_friction = d(3 * rows); RB.store_twist = d(3 * rows);
  RB.store_slip = d(3 * rows); RB.mu = d(rows); RB.e = d(rows);
  Work& W = T.W;
  W.n = work;
  W.base = d((size_t)Work::WIDTH * work);
  W.sliding = reinterpret_cast<uint8_t*>(take(work));
  WorkBar& WB = T.WB;
  WB.n = work;
  WB.base = d((size_t)WorkBar::WIDTH * work);
  B.branch = nullptr;
  R.branch = nullptr;
  T.body_branch = reinterpret_cast<uint64_t*>(take(nb * 8));
  T.row_branch = reinterpret_cast<uint64_t*>(take((rows ? rows : 1) * 8));
  T.checkpoints = d((size_t)(h.steps / h.every + 1) * nb * STATE);
  T.row_tape = d((size_t)2 * most_sub * work * ROW_TAPE);
  T.body_tape = d((size_t)h.every * 2 * most_sub * nb * STATE);
  return (at + 255) & ~(size_t)255;
}

// Each world's branch record from its bodies' and rows' (held_forward's note()): every record
// mixed with its body's or row's place in the world, then summed, so the fold does not
// depend on the order the records are read in.
inline void fold_branches(int worlds, int steps, const int32_t* body_offset, const int32_t* row_offset,
                          const uint64_t* body, const uint64_t* row, uint64_t* out) {
  auto mix = [](uint64_t where, uint64_t v) {
    uint64_t x = v ^ (where * 0x9e3779b97f4a7c15ull);
    x = (x ^ (x >> 31)) * 0xbf58476d1ce4e5b9ull;
    return x ^ (x >> 29);
  };
  for (int w = 0; w < worlds; w++) {
    uint64_t acc = 0;
    for (int g = body_offset[w]; g < body_offset[w + 1]; g++) acc += mix((uint64_t)(g - body_offset[w]), body[g]);
    for (int t = 0; t < steps; t++) {
      const int r0 = row_offset[t * worlds + w], r1 = row_offset[t * worlds + w + 1];
      for (int r = r0; r < r1; r++) acc += mix(((uint64_t)(t + 1) << 32) | (uint64_t)(r - r0), row[r]);
    }
    out[w] = acc;
  }
}

// A world's bodies at the start of the forward pass: p q v w from state0, the tumble's
// state as recorded, the mass columns from the mass.
HD void start_world(const Trajectory& T, const HeldIO& io, int wld, int tid, int threads) {
  const Bodies& B = T.B;
  for (int g = T.body_offset[wld] + tid; g < T.body_offset[wld + 1]; g += threads) {
    const double* s = io.state0 + (size_t)BM_HELD_STATE * g;
    for (int k = 0; k < 3; k++) { B.p[3 * g + k] = s[k]; B.v[3 * g + k] = s[7 + k]; B.w[3 * g + k] = s[10 + k]; }
    for (int k = 0; k < 4; k++) B.q[4 * g + k] = s[3 + k];
    for (int k = 0; k < 3; k++) B.left[3 * g + k] = io.left0[3 * g + k];
    B.ceiling[g] = io.ceiling0[g];
    from_mass(B, g);
  }
}

HD void end_world(const Trajectory& T, const HeldIO& io, int wld, int tid, int threads) {
  const Bodies& B = T.B;
  for (int g = T.body_offset[wld] + tid; g < T.body_offset[wld + 1]; g += threads) {
    double* s = io.state_out + (size_t)BM_HELD_STATE * g;
    for (int k = 0; k < 3; k++) { s[k] = B.p[3 * g + k]; s[7 + k] = B.v[3 * g + k]; s[10 + k] = B.w[3 * g + k]; }
    for (int k = 0; k < 4; k++) s[3 + k] = B.q[4 * g + k];
  }
}

// The backward pass's start: the final state's adjoints, every other adjoint 0.
HD void start_bars(const Trajectory& T, const HeldIO& io, int wld, int tid, int threads) {
  const BodiesBar& BB = T.BB;
  for (int g = T.body_offset[wld] + tid; g < T.body_offset[wld + 1]; g += threads) {
    const double* s = io.state_bar + (size_t)BM_HELD_STATE * g;
    for (int k = 0; k < 3; k++) {
      BB.p[3 * g + k] = s[k]; BB.v[3 * g + k] = s[7 + k]; BB.w[3 * g + k] = s[10 + k]; BB.left[3 * g + k] = 0;
      BB.start[3 * g + k] = 0; BB.inertia[3 * g + k] = 0; BB.inv_inertia[3 * g + k] = 0;
    }
    for (int k = 0; k < 4; k++) { BB.q[4 * g + k] = s[3 + k]; BB.start_q[4 * g + k] = 0; }
    for (int k = 0; k < 6; k++) BB.inv_i[6 * g + k] = 0;
    BB.ceiling[g] = 0;
    BB.inv_mass[g] = 0;
  }
}

HD void end_bars(const Trajectory& T, const HeldIO& io, int wld, int tid, int threads) {
  const BodiesBar& BB = T.BB;
  for (int g = T.body_offset[wld] + tid; g < T.body_offset[wld + 1]; g += threads) {
    double* s = io.state0_bar + (size_t)BM_HELD_STATE * g;
    for (int k = 0; k < 3; k++) { s[k] = BB.p[3 * g + k]; s[7 + k] = BB.v[3 * g + k]; s[10 + k] = BB.w[3 * g + k]; }
    for (int k = 0; k < 4; k++) s[3 + k] = BB.q[4 * g + k];
  }
}

}  // namespace adj
}  // namespace bm
