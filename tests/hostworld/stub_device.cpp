// TEST ONLY. The library's device calls (bm_upload, bm_colour, bm_soft_step,
// bm_download, bm_collide_host) answered on the CPU by the oracle, so the host half of
// bm_world (bearing_amd/csrc/world.cpp: trees, pair list, narrow order, freezing and
// sleeping) can be run against the reference's trajectories on a machine with no GPU.
// The product library never links this: there each of these calls runs on the device.
#include <cstring>
#include <string>
#include <vector>

#include "../../bearing_amd/csrc/broad_dev.h"
#include "../../bearing_amd/csrc/held_layout.h"
#include "../../include/bearing_mi355x.h"
#include "../../oracle/oracle.h"

extern "C" int oracle_collide(int n, const double* a, const double* b, const double* margin, const double* tolerance,
                              int32_t* count, double* normal, double* points, double* separations, uint32_t* ids);
extern "C" void oracle_losses(double* out);
extern "C" int oracle_meet(int n, const double* a, const double* b, const double* pair, int32_t* count,
                           double* normal, double* points, double* separations, uint32_t* ids, double* shift,
                           int32_t* flags, double* loss);
extern "C" void oracle_fields(int n, const int32_t* columns, const int32_t* rows, const double* spacing,
                              const int64_t* off, const double* heights);
extern "C" int oracle_level(int n, const int32_t* field, const double* position, const double* axes,
                            const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                            int cap, int32_t* manifolds, int32_t* count, double* normal, double* points,
                            double* separations, uint32_t* ids);
extern "C" void oracle_hulls(int n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l,
                             const double* vertices, const int32_t* start, const int32_t* loops, const double* planes);

// The held trajectory (bm_held_*), stepped on the host by the library's own adjoint code
// (bearing_amd/csrc/adjoint.h), one world after another, as one thread.
struct HeldHost {
  std::vector<char> mem;
  bm::adj::Trajectory T;
  bm::adj::HeldIO io;
};

// The broadphase passes' device memory, as copies: what bm_broad_set_bodies, _set_state, _keep
// and _set_trees gave, so a body the host changed without saying so keeps its old state here.
struct BroadHost {
  int n = 0;
  std::vector<uint8_t> type, shape;
  std::vector<double> dims, gs, reach, size, local, position, rotation, velocity, angular;
  struct T {
    int root = -1;
    std::vector<bm::broad::Node> node;
  } trees[2];
  std::vector<int> leaf;
  std::vector<int8_t> which;
  bool have_trees = false;
  bm::broad::Bodies bodies() const {
    return bm::broad::Bodies{type.data(), shape.data(), dims.data(), gs.data(), reach.data(), size.data(),
                             local.data(), position.data(), rotation.data(), velocity.data(), angular.data()};
  }
  bm::broad::Tree tree(int t) const { return bm::broad::Tree{trees[t].root, trees[t].node.data()}; }
};

struct bm_context {
  std::vector<bm_columns> hs;
  std::string error;
  HeldHost* held = nullptr;
  BroadHost broad;
};

namespace {
// The columns narrow() writes, moved to their rows by colour() (src/graph.js:178-214).
template <class T>
void arrange(T* col, int per, const int32_t* order, int n) {
  std::vector<T> from(col, col + (size_t)per * n);
  for (int o = 0; o < n; o++)
    for (int k = 0; k < per; k++) col[(size_t)o * per + k] = from[(size_t)order[o] * per + k];
}
template <class T>
T* mut(const T* p) { return const_cast<T*>(p); }
}  // namespace

extern "C" {

int bm_create(bm_context** out, int32_t) {
  *out = new bm_context();
  return BM_OK;
}
int bm_destroy(bm_context* ctx) {
  if (ctx) delete ctx->held;
  delete ctx;
  return BM_OK;
}

int bm_held_load(bm_context* ctx, const bm_held* h) {
  using namespace bm::adj;
  delete ctx->held;
  ctx->held = new HeldHost();
  HeldHost& H = *ctx->held;
  std::vector<int32_t> work_at;
  const size_t bytes = lay_out(*h, nullptr, H.T, H.io, nullptr, work_at);
  H.mem.assign(bytes + 256, 0);
  char* base = reinterpret_cast<char*>(((uintptr_t)H.mem.data() + 255) & ~(uintptr_t)255);
  std::vector<Copy> copies;
  lay_out(*h, base, H.T, H.io, &copies, work_at);
  for (const Copy& c : copies) std::memcpy(c.to, c.from, c.bytes);
  return BM_OK;
}

int bm_held_forward(bm_context* ctx, const double* state0, const double* mass, const double* friction,
                    const double* restitution, double* state_out) {
  using namespace bm::adj;
  if (!ctx->held) return BM_ERR_ARGUMENT;
  HeldHost& H = *ctx->held;
  const int nb = H.T.bodies;
  std::memcpy(H.io.state0, state0, sizeof(double) * BM_HELD_STATE * nb);
  std::memcpy(H.io.mass_in, mass, sizeof(double) * nb);
  std::memcpy(H.io.friction_in, friction, sizeof(double) * nb);
  std::memcpy(H.io.restitution_in, restitution, sizeof(double) * nb);
  const Ctx c{0, 1, nullptr};
  for (int w = 0; w < H.T.worlds; w++) {
    start_world(H.T, H.io, w, 0, 1);
    forward_world(c, [] {}, H.T, w);
    end_world(H.T, H.io, w, 0, 1);
  }
  std::memcpy(state_out, H.io.state_out, sizeof(double) * BM_HELD_STATE * nb);
  return BM_OK;
}

int bm_held_backward(bm_context* ctx, const double* state_bar, double* state0_bar, double* mass_bar,
                     double* friction_bar, double* restitution_bar) {
  using namespace bm::adj;
  if (!ctx->held) return BM_ERR_ARGUMENT;
  HeldHost& H = *ctx->held;
  const Trajectory& T = H.T;
  const int nb = T.bodies;
  const size_t rows = T.row_offset[T.steps * T.worlds];
  std::memcpy(H.io.state_bar, state_bar, sizeof(double) * BM_HELD_STATE * nb);
  for (double* p : {T.RB.store_impulse}) std::memset(p, 0, sizeof(double) * P * rows);
  for (double* p : {T.RB.store_friction, T.RB.store_twist, T.RB.store_slip}) std::memset(p, 0, sizeof(double) * 3 * rows);
  for (double* p : {T.RB.mu, T.RB.e}) std::memset(p, 0, sizeof(double) * rows);
  double slot[ROW_SLOT];
  const Ctx c{0, 1, slot};
  for (int w = 0; w < T.worlds; w++) {
    start_bars(T, H.io, w, 0, 1);
    backward_world(c, [] {}, T, w);
    end_bars(T, H.io, w, 0, 1);
    parameters_world(T, w, H.io.mass_bar, H.io.friction_bar, H.io.restitution_bar);
  }
  std::memcpy(state0_bar, H.io.state0_bar, sizeof(double) * BM_HELD_STATE * nb);
  std::memcpy(mass_bar, H.io.mass_bar, sizeof(double) * nb);
  std::memcpy(friction_bar, H.io.friction_bar, sizeof(double) * nb);
  std::memcpy(restitution_bar, H.io.restitution_bar, sizeof(double) * nb);
  return BM_OK;
}

int bm_held_branches(bm_context* ctx, uint64_t* per_world) {
  if (!ctx->held) return BM_ERR_ARGUMENT;
  const bm::adj::Trajectory& T = ctx->held->T;
  bm::adj::fold_branches(T.worlds, T.steps, T.body_offset, T.row_offset, T.body_branch, T.row_branch, per_world);
  return BM_OK;
}

int bm_held_times(const bm_context*, double* forward_ms, double* backward_ms) {
  if (forward_ms) *forward_ms = 0;
  if (backward_ms) *backward_ms = 0;
  return BM_OK;
}
const char* bm_last_error(const bm_context* ctx) { return ctx->error.c_str(); }
double bm_last_kernel_ms(const bm_context*) { return 0; }
int bm_upload(bm_context* ctx, const bm_columns* host) {
  if (host->n_worlds != 1) return BM_ERR_ARGUMENT;
  ctx->hs.assign(1, *host);
  return BM_OK;
}
int bm_download(bm_context*, bm_columns*) { return BM_OK; }
// The worlds' own arrays are worked on in place, so nothing moves either way.
int bm_upload_worlds(bm_context* ctx, int32_t n, const bm_columns* const* worlds) {
  ctx->hs.clear();
  for (int i = 0; i < n; i++) {
    if (worlds[i]->n_worlds != 1) return BM_ERR_ARGUMENT;
    ctx->hs.push_back(*worlds[i]);
  }
  return BM_OK;
}
int bm_download_worlds(bm_context*, int32_t, bm_columns* const*) { return BM_OK; }
int bm_set_device_only(bm_context*, int32_t) { return BM_OK; }
int bm_set_body_version(bm_context*, uint64_t) { return BM_OK; }
int bm_set_graph(bm_context*, int32_t) { return BM_OK; }
int bm_graph_counts(const bm_context*, int64_t* out) {
  out[0] = out[1] = out[2] = 0;
  return BM_OK;
}

}  // extern "C"

namespace {
int colour_one(bm_columns& h) {
  const int n = h.manifold_offset[1], top = h.body_offset[1];
  int frozen = 0;
  for (int s = 0; s < top; s++) frozen += h.frozen[s] != 0;
  oracle_colour(n, h.a, h.b, top, h.type, h.frozen, frozen, h.colour, h.order, h.row, h.starts, h.solved);
  const int32_t* order = h.order;
  arrange(h.a, 1, order, n); arrange(h.b, 1, order, n); arrange(h.pair, 1, order, n); arrange(h.sub, 1, order, n);
  arrange(h.points, 1, order, n); arrange(h.colour, 1, order, n); arrange(h.pid, 4, order, n);
  arrange(h.normal, 3, order, n); arrange(h.t1, 3, order, n); arrange(h.t2, 3, order, n);
  arrange(h.friction, 1, order, n); arrange(h.restitution, 1, order, n); arrange(h.tolerance, 1, order, n);
  arrange(h.patch, 1, order, n); arrange(h.roll_impulse1, 1, order, n); arrange(h.roll_impulse2, 1, order, n);
  arrange(h.spin_impulse, 1, order, n); arrange(h.tangent_impulse1, 1, order, n);
  arrange(h.tangent_impulse2, 1, order, n); arrange(h.drift1, 1, order, n); arrange(h.drift2, 1, order, n);
  arrange(h.separation, 4, order, n); arrange(h.normal_impulse, 4, order, n); arrange(h.point, 12, order, n);
  arrange(h.shift, 12, order, n);
  // rest() (src/contact.js:1318-1329)
  for (int m = *h.solved; m < n; m++) {
    double least = __builtin_inf();
    for (int k = 0; k < h.points[m]; k++) {
      const int i = 4 * m + k;
      if (h.separation[i] < least) least = h.separation[i];
      h.max_normal_impulse[i] = h.normal_impulse[i];
      h.relative_velocity[i] = 0;
    }
    h.gap[m] = least;
  }
  return BM_OK;
}
}  // namespace

extern "C" {
int bm_colour(bm_context* ctx) {
  for (bm_columns& h : ctx->hs) colour_one(h);
  return BM_OK;
}
}  // extern "C"

// More threads for the oracle's soft step (oracle_soft_step_workers), 0 for none.
static int soft_workers = 0;

namespace {
int soft_one(bm_context* ctx, bm_columns& h) {
  OracleSoft o{};
  o.top = h.body_offset[1];
  o.awake_count = h.awake_offset[1];
  o.substeps = *h.substeps;
  o.count = h.manifold_offset[1];
  o.solved = *h.solved;
  o.pair_count = h.pair_offset[1];
  o.dt = *h.dt;
  for (int k = 0; k < 3; k++) o.gravity[k] = h.gravity[k];
  o.position = h.position; o.rotation = h.rotation; o.velocity = h.velocity; o.angular = h.angular;
  o.dims = mut(h.dims); o.inv_mass = mut(h.inv_mass); o.inertia = mut(h.inertia); o.inv_inertia = mut(h.inv_inertia);
  o.inv_i = h.inv_i; o.axes = mut(h.axes); o.tumbled = h.tumbled; o.ceiling = h.ceiling;
  o.gravity_scale = mut(h.gravity_scale); o.damping = mut(h.damping); o.angular_damping = mut(h.angular_damping);
  o.start = h.start; o.start_q = h.start_q;
  o.type = mut(h.type); o.shape = mut(h.shape); o.frozen = mut(h.frozen); o.lock = mut(h.lock);
  o.framed = mut(h.framed); o.awake = mut(h.awake); o.joint_head = mut(h.joint_head);
  o.awake_list = mut(h.awake_list);
  o.starts = h.starts; o.a = h.a; o.b = h.b; o.pair = h.pair; o.sub = h.sub; o.row = h.row;
  o.points = h.points; o.sliding = h.sliding;
  o.normal = h.normal; o.t1 = h.t1; o.t2 = h.t2; o.friction = h.friction; o.restitution = h.restitution;
  o.tolerance = h.tolerance; o.patch = h.patch; o.roll_impulse1 = h.roll_impulse1; o.roll_impulse2 = h.roll_impulse2;
  o.spin_impulse = h.spin_impulse; o.tangent_impulse1 = h.tangent_impulse1; o.tangent_impulse2 = h.tangent_impulse2;
  o.drift1 = h.drift1; o.drift2 = h.drift2; o.separation = h.separation; o.point = h.point; o.shift = h.shift;
  o.normal_impulse = h.normal_impulse; o.max_normal_impulse = h.max_normal_impulse;
  o.relative_velocity = h.relative_velocity; o.gap = h.gap; o.now = h.now; o.ra = h.ra; o.rb = h.rb; o.la = h.la;
  o.lb = h.lb; o.adjusted = h.adjusted; o.normal_mass = h.normal_mass; o.jac = h.jac; o.block = h.block;
  o.lever = h.lever; o.anchor_a = h.anchor_a; o.anchor_b = h.anchor_b; o.fjac = h.fjac;
  o.tangent_mass = h.tangent_mass; o.roll_mass1 = h.roll_mass1; o.roll_mass2 = h.roll_mass2;
  o.spin_mass = h.spin_mass; o.pid = h.pid;
  o.pair_points = h.pair_points; o.pair_pid = h.pair_pid; o.pair_normal_impulse = h.pair_normal_impulse;
  o.pair_twist = h.pair_twist; o.pair_friction = h.pair_friction; o.pair_slip = h.pair_slip;
  if (oracle_soft_step_workers(&o, soft_workers) != 0) {
    ctx->error = "oracle soft step: a manifold it does not cover";
    return BM_ERR_UNSUPPORTED;
  }
  return BM_OK;
}
}  // namespace

// For tests of failure handling: the next soft step of any context fails as a device error
// would, after the narrowphase has moved the world's pairs on (the step is then half done).
static int fail_next_soft_step = 0;

extern "C" {
void stub_fail_next_soft_step() { fail_next_soft_step = 1; }
void stub_soft_workers(int workers) { soft_workers = workers > 0 ? workers : 0; }

int bm_soft_step(bm_context* ctx) {
  if (fail_next_soft_step) {
    fail_next_soft_step = 0;
    ctx->error = "bm_soft_step: a device failure the test asked for";
    return BM_ERR_DEVICE;
  }
  for (bm_columns& h : ctx->hs)
    if (int rc = soft_one(ctx, h)) return rc;
  return BM_OK;
}

int bm_contact_phase(bm_context* ctx) {
  if (int rc = bm_colour(ctx)) return rc;
  return bm_soft_step(ctx);
}

int bm_collide_host(bm_context*, int32_t n, const double* a, const double* b, const double* margin,
                    const double* tolerance, int32_t* count, double* normal, double* points, double* separations,
                    uint32_t* ids, double* loss) {
  oracle_collide(n, a, b, margin, tolerance, count, normal, points, separations, ids);
  if (loss) oracle_losses(loss);
  return BM_OK;
}

int bm_set_fields(bm_context*, int32_t n, const int32_t* columns, const int32_t* rows, const double* spacing,
                  const int64_t* offsets, const double* heights) {
  oracle_fields(n, columns, rows, spacing, offsets, heights);
  return BM_OK;
}

// As the device does, every direction is counted even past cap; only cap are written.
int bm_level_host(bm_context*, int32_t n, const int32_t* field, const double* position, const double* axes,
                  const double* shapes, const double* boxes, const double* margin, const double* tolerance, int32_t cap,
                  int32_t* manifolds, int32_t* count, double* normal, double* points, double* separations,
                  uint32_t* ids) {
  const int big = 256;
  std::vector<int32_t> m(n), c((size_t)big * n);
  std::vector<double> nr(3 * (size_t)big * n), pt(12 * (size_t)big * n), sp(4 * (size_t)big * n);
  std::vector<uint32_t> id(4 * (size_t)big * n);
  if (oracle_level(n, field, position, axes, shapes, boxes, margin, tolerance, big, m.data(), c.data(), nr.data(),
                   pt.data(), sp.data(), id.data()) != n)
    return BM_ERR_UNSUPPORTED;
  for (int k = 0; k < n; k++) {
    manifolds[k] = m[k];
    for (int j = 0; j < cap; j++) {
      const size_t o = (size_t)k * cap + j, q = (size_t)k * big + j;
      const bool has = j < m[k];
      count[o] = has ? c[q] : 0;
      for (int i = 0; i < 3; i++) normal[3 * o + i] = has ? nr[3 * q + i] : 0;
      for (int i = 0; i < 12; i++) points[12 * o + i] = has ? pt[12 * q + i] : 0;
      for (int i = 0; i < 4; i++) separations[4 * o + i] = has ? sp[4 * q + i] : 0;
      for (int i = 0; i < 4; i++) ids[4 * o + i] = has ? id[4 * q + i] : 0;
    }
  }
  return BM_OK;
}

int bm_level_packed(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                    const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                    int32_t cap, int32_t* manifolds, int32_t* start, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids) {
  const size_t slots = (size_t)n * cap;
  std::vector<int32_t> c(slots);
  std::vector<double> nr(3 * slots), pt(12 * slots), sp(4 * slots);
  std::vector<uint32_t> id(4 * slots);
  if (int rc = bm_level_host(ctx, n, field, position, axes, shapes, boxes, margin, tolerance, cap, manifolds, c.data(),
                             nr.data(), pt.data(), sp.data(), id.data()))
    return rc;
  int at = 0;
  for (int k = 0; k < n; k++) {
    start[k] = at;
    const int got = manifolds[k] < cap ? manifolds[k] : cap;
    for (int j = 0; j < got; j++, at++) {
      const size_t o = (size_t)k * cap + j;
      count[at] = c[o];
      for (int i = 0; i < 3; i++) normal[3 * (size_t)at + i] = nr[3 * o + i];
      for (int i = 0; i < 12; i++) points[12 * (size_t)at + i] = pt[12 * o + i];
      for (int i = 0; i < 4; i++) separations[4 * (size_t)at + i] = sp[4 * o + i];
      for (int i = 0; i < 4; i++) ids[4 * (size_t)at + i] = id[4 * o + i];
    }
  }
  start[n] = at;
  return BM_OK;
}

int bm_set_hulls(bm_context*, int32_t n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l,
                 const double* vertices, const int32_t* start, const int32_t* loops, const double* planes) {
  oracle_hulls(n, hull_v, hull_f, hull_l, vertices, start, loops, planes);
  return BM_OK;
}

int bm_meet_host(bm_context* ctx, int32_t n, const double* a, const double* b, const double* pair, int32_t* count,
                 double* normal, double* points, double* separations, uint32_t* ids, double* shift, int32_t* flags,
                 double* loss) {
  if (oracle_meet(n, a, b, pair, count, normal, points, separations, ids, shift, flags, loss) != n) {
    ctx->error = "oracle meet: a pair of kinds it does not cover";
    return BM_ERR_UNSUPPORTED;
  }
  return BM_OK;
}

// The launch shape has no meaning on the CPU; its arguments are checked as the device's are.
int bm_set_launch(bm_context* ctx, int32_t shape, int32_t resident) {
  (void)resident;
  if (shape < 0 || shape > 2) {
    ctx->error = "bm_set_launch: shape must be 0, 1 or 2";
    return BM_ERR_ARGUMENT;
  }
  return BM_OK;
}

int bm_meet_indexed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                    const int32_t* ib, const double* pair, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids, double* shift, int32_t* flags, double* loss) {
  std::vector<double> a(32 * (size_t)n), b(32 * (size_t)n);
  for (int32_t k = 0; k < n; k++) {
    if (ia[k] < 0 || ia[k] >= bodies || ib[k] < 0 || ib[k] >= bodies) {
      ctx->error = "bm_meet_indexed: a pair names a body the table lacks";
      return BM_ERR_ARGUMENT;
    }
    std::memcpy(&a[32 * (size_t)k], table + 32 * (size_t)ia[k], 32 * sizeof(double));
    std::memcpy(&b[32 * (size_t)k], table + 32 * (size_t)ib[k], 32 * sizeof(double));
  }
  return bm_meet_host(ctx, n, a.data(), b.data(), pair, count, normal, points, separations, ids, shift, flags, loss);
}

// bm_meet_indexed's results packed as bm_meet_packed leaves them: the pairs that found a
// manifold in pair order, each one row.
int bm_meet_packed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                   const int32_t* ib, const double* pair, int32_t* count, int32_t* flags, int32_t* row,
                   double* normal, double* points, double* separations, uint32_t* ids, double* shift, double* loss) {
  const size_t N = (size_t)n;
  std::vector<double> nm(3 * N), pt(12 * N), sp(4 * N), sh(12 * N), ls(N);
  std::vector<uint32_t> id(4 * N);
  if (int rc = bm_meet_indexed(ctx, bodies, table, n, ia, ib, pair, count, nm.data(), pt.data(), sp.data(), id.data(),
                               sh.data(), flags, ls.data()))
    return rc;
  int32_t r = 0;
  for (size_t k = 0; k < N; k++) {
    if (count[k] <= 0) {
      row[k] = -1;
      continue;
    }
    row[k] = r;
    std::memcpy(normal + 3 * (size_t)r, &nm[3 * k], 3 * sizeof(double));
    std::memcpy(points + 12 * (size_t)r, &pt[12 * k], 12 * sizeof(double));
    std::memcpy(separations + 4 * (size_t)r, &sp[4 * k], 4 * sizeof(double));
    std::memcpy(shift + 12 * (size_t)r, &sh[12 * k], 12 * sizeof(double));
    std::memcpy(ids + 4 * (size_t)r, &id[4 * k], 4 * sizeof(uint32_t));
    loss[r] = ls[k];
    r++;
  }
  return BM_OK;
}

}  // extern "C"

// The broadphase passes (bm_broad_*) on the CPU, with the library's own per-body and per-pair
// arithmetic (bearing_amd/csrc/broad_dev.h) on the copies above. broad_calls counts the rows
// each pass was given (bounds, walks, holds), for tests to see that the passes ran.
static int64_t broad_calls[3] = {0, 0, 0};
extern "C" {
int64_t stub_broad_rows(int pass) { return pass >= 0 && pass < 3 ? broad_calls[pass] : -1; }

int bm_broad_set_bodies(bm_context* ctx, const bm_broad_bodies* b) {
  BroadHost& B = ctx->broad;
  const size_t n = b->n;
  B.n = b->n;
  B.type.assign(b->type, b->type + n);
  B.shape.assign(b->shape, b->shape + n);
  B.dims.assign(b->dims, b->dims + 3 * n);
  B.gs.assign(b->gravity_scale, b->gravity_scale + n);
  B.reach.assign(b->reach, b->reach + n);
  B.size.assign(b->size, b->size + n);
  B.local.assign(b->local, b->local + 6 * n);
  B.position.assign(b->position, b->position + 3 * n);
  B.rotation.assign(b->rotation, b->rotation + 4 * n);
  B.velocity.assign(b->velocity, b->velocity + 3 * n);
  B.angular.assign(b->angular, b->angular + 3 * n);
  B.have_trees = false;
  return BM_OK;
}

int bm_broad_set_state(bm_context* ctx, int32_t n, const int32_t* slots, const double* position,
                       const double* rotation, const double* velocity, const double* angular) {
  BroadHost& B = ctx->broad;
  for (int32_t i = 0; i < n; i++) {
    const size_t s = slots[i];
    if (slots[i] < 0 || slots[i] >= B.n) {
      ctx->error = "bm_broad_set_state: a slot is not a body bm_broad_set_bodies gave";
      return BM_ERR_ARGUMENT;
    }
    for (int k = 0; k < 3; k++) B.position[3 * s + k] = position[3 * s + k];
    for (int k = 0; k < 4; k++) B.rotation[4 * s + k] = rotation[4 * s + k];
    for (int k = 0; k < 3; k++) B.velocity[3 * s + k] = velocity[3 * s + k];
    for (int k = 0; k < 3; k++) B.angular[3 * s + k] = angular[3 * s + k];
  }
  return BM_OK;
}

int bm_broad_keep(bm_context* ctx) {
  BroadHost& B = ctx->broad;
  if (ctx->hs.size() != 1 || ctx->hs[0].body_offset[1] != B.n) {
    ctx->error = "bm_broad_keep: the last batch is not one world of the broadphase's bodies";
    return BM_ERR_ARGUMENT;
  }
  const bm_columns& h = ctx->hs[0];
  const size_t n = B.n;
  B.position.assign(h.position, h.position + 3 * n);
  B.rotation.assign(h.rotation, h.rotation + 4 * n);
  B.velocity.assign(h.velocity, h.velocity + 3 * n);
  B.angular.assign(h.angular, h.angular + 3 * n);
  return BM_OK;
}

int bm_broad_set_trees(bm_context* ctx, const bm_broad_tree* trees, int32_t n, const int32_t* leaf,
                       const int8_t* which) {
  BroadHost& B = ctx->broad;
  if (n != B.n) {
    ctx->error = "bm_broad_set_trees: the trees need every body bm_broad_set_bodies gave";
    return BM_ERR_ARGUMENT;
  }
  for (int t = 0; t < 2; t++) {
    const bm_broad_tree& x = trees[t];
    const size_t m = x.nodes;
    BroadHost::T& y = B.trees[t];
    y.root = x.root;
    y.node.resize(m);
    for (size_t i = 0; i < m; i++) {
      for (int k = 0; k < 6; k++) y.node[i].box[k] = x.box[6 * i + k];
      y.node[i].child1 = x.child1[i]; y.node[i].child2 = x.child2[i];
      y.node[i].height = x.height[i]; y.node[i].item = x.item[i];
    }
  }
  B.leaf.assign(leaf, leaf + n);
  B.which.assign(which, which + n);
  B.have_trees = true;
  return BM_OK;
}

int bm_broad_bounds(bm_context* ctx, int32_t n, const int32_t* list, double dt, double gravity_length, double* out,
                    uint8_t* refile) {
  const BroadHost& B = ctx->broad;
  if (!B.have_trees) {
    ctx->error = "bm_broad_bounds: no trees (bm_broad_set_trees)";
    return BM_ERR_ARGUMENT;
  }
  const bm::broad::Bodies bodies = B.bodies();
  const bm::broad::Tree moving = B.tree(1);
  broad_calls[0] += n;
  for (int32_t i = 0; i < n; i++) {
    const int s = list[i];
    double o[16];
    bm::broad::bounds(bodies, s, dt, gravity_length, o);
    for (int k = 0; k < 16; k++) out[(size_t)k * n + i] = o[k];
    refile[i] = B.which[s] == bm::broad::MOVING_TREE && !bm::broad::contains(moving.node[B.leaf[s]].box, o + 10);
  }
  return BM_OK;
}

int bm_broad_query(bm_context* ctx, int32_t n, const int32_t* moved, int32_t* counts, int32_t* starts, int32_t* hits,
                   int64_t capacity, int64_t* total) {
  const BroadHost& B = ctx->broad;
  std::vector<int32_t> all;
  broad_calls[1] += n;
  for (int32_t i = 0; i < n; i++) {
    const int a = moved[i], t = B.which[a];
    starts[i] = (int32_t)all.size();
    counts[i] = 0;
    if (t < 0) continue;
    double q[6];
    for (int k = 0; k < 6; k++) q[k] = B.trees[t].node[B.leaf[a]].box[k];
    auto put = [&](int b) {
      if (b != a) all.push_back(b);
    };
    bool ok = true;
    int stack[bm::broad::STACK];
    if (B.type[a] != bm::broad::STATIC) ok = bm::broad::query(B.tree(0), q, stack, 1, put);
    if (ok) ok = bm::broad::query(B.tree(1), q, stack, 1, put);
    if (!ok) {
      ctx->error = "bm_broad_query: a tree is deeper than a walk's stack holds";
      return BM_ERR_DEVICE;
    }
    counts[i] = (int32_t)all.size() - starts[i];
  }
  *total = (int64_t)all.size();
  if (*total <= capacity && !all.empty()) std::memcpy(hits, all.data(), 4 * all.size());
  return BM_OK;
}

int bm_broad_holds(bm_context* ctx, int32_t n, const int32_t* a, const int32_t* b, uint8_t* apart) {
  const BroadHost& B = ctx->broad;
  broad_calls[2] += n;
  for (int32_t i = 0; i < n; i++) {
    const int ta = B.which[a[i]], tb = B.which[b[i]];
    if (ta < 0 || tb < 0) {
      apart[i] = 0;
      continue;
    }
    apart[i] = !bm::broad::overlap(B.trees[ta].node[B.leaf[a[i]]].box, B.trees[tb].node[B.leaf[b[i]]].box);
  }
  return BM_OK;
}

}  // extern "C"
