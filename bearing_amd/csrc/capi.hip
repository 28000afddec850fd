// The C ABI of include/bearing_mi355x.h: contexts, HBM residency of a batch's
// columns, and the soft-step launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/bearing_mi355x.h"
#include "adjoint_dev.h"
#include "broad.h"
#include "collide.h"
#include "colour.h"
#include "soft.h"

using bm::SoftDev;

namespace {

// KEPT: what the pair list keeps, by pair on the host; on the device by manifold row (each
// solved row stores its pair's memory there), never uploaded, scattered to the pairs on download.
enum Unit { WORLD1, WORLD, STARTS, WORLD3, BODY, AWAKE, MANIFOLD, KEPT };
enum Dir { IN = 1, IO = 3 };

struct Column {
  const char* name;
  size_t host;  // offsetof in bm_columns
  size_t dev;   // offsetof in SoftDev
  size_t elem;  // bytes per number
  int per;      // numbers per unit
  Unit unit;
  Dir dir;
};

#define COL(n, e, p, u, d) {#n, offsetof(bm_columns, n), offsetof(SoftDev, n), e, p, u, d}
const Column COLUMNS[] = {
    COL(body_offset, 4, 1, WORLD1, IN), COL(manifold_offset, 4, 1, WORLD1, IN), COL(pair_offset, 4, 1, WORLD1, IN),
    COL(awake_offset, 4, 1, WORLD1, IN), COL(starts, 4, 1, STARTS, IO), COL(solved, 4, 1, WORLD, IO),
    COL(substeps, 4, 1, WORLD, IN), COL(dt, 8, 1, WORLD, IN), COL(gravity, 8, 1, WORLD3, IN),
    COL(position, 8, 3, BODY, IO), COL(rotation, 8, 4, BODY, IO), COL(velocity, 8, 3, BODY, IO),
    COL(angular, 8, 3, BODY, IO), COL(start, 8, 3, BODY, IO), COL(start_q, 8, 4, BODY, IO),
    COL(inv_i, 8, 6, BODY, IO), COL(tumbled, 8, 3, BODY, IO), COL(ceiling, 8, 1, BODY, IO),
    COL(dims, 8, 3, BODY, IN), COL(inv_mass, 8, 1, BODY, IN), COL(inertia, 8, 3, BODY, IN),
    COL(inv_inertia, 8, 3, BODY, IN), COL(axes, 8, 9, BODY, IN), COL(gravity_scale, 8, 1, BODY, IN),
    COL(damping, 8, 1, BODY, IN), COL(angular_damping, 8, 1, BODY, IN), COL(type, 1, 1, BODY, IN),
    COL(shape, 1, 1, BODY, IN), COL(frozen, 1, 1, BODY, IN), COL(lock, 1, 1, BODY, IN), COL(framed, 1, 1, BODY, IN),
    COL(awake, 1, 1, BODY, IN), COL(joint_head, 4, 1, BODY, IN), COL(awake_list, 4, 1, AWAKE, IN),
    COL(a, 4, 1, MANIFOLD, IO), COL(b, 4, 1, MANIFOLD, IO), COL(pair, 4, 1, MANIFOLD, IO),
    COL(sub, 4, 1, MANIFOLD, IO), COL(row, 4, 1, MANIFOLD, IO), COL(colour, 4, 1, MANIFOLD, IO),
    COL(order, 4, 1, MANIFOLD, IO), COL(points, 1, 1, MANIFOLD, IO),
    COL(sliding, 1, 1, MANIFOLD, IO), COL(normal, 8, 3, MANIFOLD, IO), COL(t1, 8, 3, MANIFOLD, IO),
    COL(t2, 8, 3, MANIFOLD, IO), COL(friction, 8, 1, MANIFOLD, IO), COL(restitution, 8, 1, MANIFOLD, IO),
    COL(tolerance, 8, 1, MANIFOLD, IO), COL(patch, 8, 1, MANIFOLD, IO), COL(roll_impulse1, 8, 1, MANIFOLD, IO),
    COL(roll_impulse2, 8, 1, MANIFOLD, IO), COL(spin_impulse, 8, 1, MANIFOLD, IO),
    COL(tangent_impulse1, 8, 1, MANIFOLD, IO), COL(tangent_impulse2, 8, 1, MANIFOLD, IO),
    COL(drift1, 8, 1, MANIFOLD, IO), COL(drift2, 8, 1, MANIFOLD, IO), COL(separation, 8, 4, MANIFOLD, IO),
    COL(point, 8, 12, MANIFOLD, IO), COL(shift, 8, 12, MANIFOLD, IO), COL(normal_impulse, 8, 4, MANIFOLD, IO),
    COL(max_normal_impulse, 8, 4, MANIFOLD, IO), COL(relative_velocity, 8, 4, MANIFOLD, IO),
    COL(gap, 8, 1, MANIFOLD, IO), COL(now, 8, 4, MANIFOLD, IO), COL(ra, 8, 12, MANIFOLD, IO),
    COL(rb, 8, 12, MANIFOLD, IO), COL(la, 8, 12, MANIFOLD, IO), COL(lb, 8, 12, MANIFOLD, IO),
    COL(adjusted, 8, 4, MANIFOLD, IO), COL(normal_mass, 8, 4, MANIFOLD, IO), COL(jac, 8, 48, MANIFOLD, IO),
    COL(block, 8, 16, MANIFOLD, IO), COL(lever, 8, 4, MANIFOLD, IO), COL(anchor_a, 8, 3, MANIFOLD, IO),
    COL(anchor_b, 8, 3, MANIFOLD, IO), COL(fjac, 8, 24, MANIFOLD, IO), COL(tangent_mass, 8, 3, MANIFOLD, IO),
    COL(roll_mass1, 8, 1, MANIFOLD, IO), COL(roll_mass2, 8, 1, MANIFOLD, IO), COL(spin_mass, 8, 1, MANIFOLD, IO),
    COL(pid, 4, 4, MANIFOLD, IO), COL(pair_points, 1, 1, KEPT, IO), COL(pair_pid, 4, 4, KEPT, IO),
    COL(pair_normal_impulse, 8, 4, KEPT, IO), COL(pair_twist, 8, 3, KEPT, IO), COL(pair_friction, 8, 3, KEPT, IO),
    COL(pair_slip, 8, 3, KEPT, IO),
};
#undef COL
constexpr int NCOL = sizeof(COLUMNS) / sizeof(COLUMNS[0]);

// The columns narrow() writes, which colour()'s arrange moves to their rows (src/graph.js:83-91).
const char* const ARRANGED[] = {"a", "b", "pair", "sub", "points", "colour", "pid", "normal", "t1", "t2",
                                "friction", "restitution", "tolerance", "patch", "roll_impulse1", "roll_impulse2",
                                "spin_impulse", "tangent_impulse1", "tangent_impulse2", "drift1", "drift2",
                                "separation", "normal_impulse", "point", "shift"};
constexpr int NARRANGED = sizeof(ARRANGED) / sizeof(ARRANGED[0]);
static_assert(NARRANGED <= bm::ARRANGE_MAX, "arrange takes at most ARRANGE_MAX columns");

int column_index(const char* name) {
  for (int c = 0; c < NCOL; c++)
    if (std::strcmp(COLUMNS[c].name, name) == 0) return c;
  return -1;
}

// Device scratch the colouring uses, grown as batches need.
struct Scratch {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct bm_context {
  int device = 0;
  hipStream_t stream = nullptr;
  void* noalloc_block = nullptr;  // this device's block for bm_level_noalloc (collide.hip)
  bool stored = false;            // a soft step ran on the batch uploaded last (its pair stores are set)
  hipEvent_t begin = nullptr, end = nullptr;
  // A download's pieces (made on first use): each piece's columns go to the worlds while the
  // pieces after it are still coming back.
  static constexpr int PIECES = 4;
  hipEvent_t piece[PIECES] = {};
  std::string error;
  int32_t n_worlds = 0, bodies = 0, awake = 0, manifolds = 0, pairs = 0;
  // Every column lives in one device arena at a 256-byte aligned offset, the columns
  // arrange() moves first, so a batch goes up and comes back in one copy each way
  // through a pinned staging buffer.
  void* buffer[NCOL] = {};   // arena + offset[c]
  size_t offset[NCOL] = {};
  size_t bytes[NCOL] = {};
  void* arena = nullptr;
  size_t arena_cap = 0, total = 0, arranged_end = 0;
  void* spare_arena = nullptr;  // arrange()'s output, copied back over the arranged block
  size_t spare_cap = 0;
  void* stage = nullptr;        // pinned host staging for uploads and downloads
  size_t stage_cap = 0;
  void* pstage = nullptr;       // pinned host staging for narrowphase batches
  size_t pstage_cap = 0;
  std::vector<int32_t> moff_host, boff_host;
  std::vector<int32_t> join_body, join_manifold, join_pair, join_awake;  // the joined offset tables
  size_t colour_lds_limit = 0;  // dynamic LDS a colouring workgroup may take (0: not yet asked)
  int32_t colour_in_lds = 1;    // colour worlds that fit in LDS there
  bm::HeldDevice* held = nullptr;  // a held trajectory (bm_held_load), for the adjoint
  SoftDev dev{};
  bool uploaded = false;
  float last_ms = 0;
  int32_t most_bodies = 0;  // the most bodies any world of the batch has
  size_t lds_limit = 0;     // dynamic LDS a workgroup may take (0: not yet asked)
  int32_t shape = 0;        // launch shape of the soft-step kernel
  int32_t resident = 1;     // keep velocities in LDS when they fit
  bool laid_out = false;    // rows are laid out by colour (checked at upload, or by bm_colour)
  Scratch deg, adj, keys, key_offset, world_of;
  // Worlds by decreasing manifold count, the order the colouring's and soft step's workgroups
  // take them in (a batch of more than one world), and its host copy.
  Scratch world_order;
  std::vector<int32_t> order_host;
  Scratch pairs_in, pairs_out;
  // The hull table (bm_set_hulls): ints (info, start, loops, edges, sides) and doubles
  // (dinfo, vertices, planes) on the device, and where each part starts.
  Scratch hull_ints, hull_doubles;
  bm::HullTable hulls{};
  int32_t hull_count = 0;
  // The heightfields (bm_set_fields): info and offsets, spacing and corner, heights; and a
  // host copy of each one's shape, for sizing touchLevel's scratch.
  Scratch field_ints, field_doubles, level_scratch;
  Scratch level_packed;  // bm_level_packed's starts and packed rows
  // A large world's colouring: sort keys and values, and the radix sort's own scratch; which
  // path the last bm_colour took (0 HBM fixed point, 1 in LDS, 2 streamed and sorted).
  Scratch colour_sort;
  Scratch wide_bar;  // the grid-wide soft step's barrier counters (bm::WIDE_WORDS), zeroed once
  Scratch wide_pick;  // the grid-wide soft step's rows grouped by path, an int a manifold
  int last_colour_path = 0;
  // A single world of at least wide_min manifolds takes the grid-wide soft step (0: never);
  // last_wide, the workgroups the last step used (0: one a world).
  int32_t wide_min = 16384, last_wide = 0;
  // Grid-wide launches the device refused (the step went the one-workgroup way), and why.
  int64_t wide_refused = 0;
  std::string wide_error;
  // The soft step's own columns stay on the device (bm_set_device_only); the arena's bytes
  // up to them, which upload and download move.
  bool keep_device_only = false;
  size_t device_only_at = 0, results_only_at = 0, moved = 0;
  // bm_contact_phase's graph (bm_set_graph): the executable graph of the colouring's and soft
  // step's launches, and how many times it was launched, made and updated.
  bool graph_on = false;
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_launches = 0, graph_instantiations = 0, graph_updates = 0;
  // The bodies' constant columns (CONSTANT), when a body version is set: their block, the
  // version and bodies it holds, and whether each column lives there for the current upload.
  uint64_t body_version = 0, kept_version = 0;
  int32_t kept_bodies = -1;
  void* constants = nullptr;
  size_t constants_cap = 0;
  bool in_block[NCOL] = {};
  bm::FieldTable fields{};
  struct FieldShape {
    int columns, rows;
    double spacing, x0, z0;
  };
  std::vector<FieldShape> field_shapes;
  std::vector<int32_t> key_offsets;
  // The broadphase passes (bm_broad_*): one world's bodies by slot (25 doubles, then type and
  // shape), its two trees (a 64-byte record a node), each body's leaf and tree, and the passes'
  // rows and results.
  struct Broad {
    int32_t n = 0;
    Scratch bodies, trees[2], leaf, rows, out, hits, counters;
    int32_t root[2] = {-1, -1}, nodes[2] = {0, 0};
    bool have_trees = false;
  } broad;
};

namespace {

int fail(bm_context* ctx, int code, const std::string& message) {
  if (ctx) ctx->error = message;
  return code;
}

#define HIP_OR_FAIL(ctx, call, where)                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(ctx, BM_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t units(const bm_context* ctx, Unit u) {
  switch (u) {
    case WORLD1: return (size_t)ctx->n_worlds + 1;
    case WORLD: return (size_t)ctx->n_worlds;
    case STARTS: return (size_t)ctx->n_worlds * BM_STARTS;
    case WORLD3: return (size_t)ctx->n_worlds * 3;
    case BODY: return (size_t)ctx->bodies;
    case AWAKE: return (size_t)ctx->awake;
    case MANIFOLD: return (size_t)ctx->manifolds;
    case KEPT: return (size_t)ctx->manifolds;
  }
  return 0;
}

template <typename T>
T field(const void* base, size_t off) {
  T v;
  std::memcpy(&v, static_cast<const char*>(base) + off, sizeof(T));
  return v;
}

// Checks a batch's offset tables and the indices inside its columns, so no kernel
// can read or write outside the buffers it is given.
std::string check_part(const bm_columns* h) {
  const int n = h->n_worlds;
  if (n <= 0) return "bm_upload: n_worlds must be 1 or more";
  const int32_t* tables[4] = {h->body_offset, h->manifold_offset, h->pair_offset, h->awake_offset};
  const char* names[4] = {"body_offset", "manifold_offset", "pair_offset", "awake_offset"};
  for (int t = 0; t < 4; t++) {
    if (!tables[t]) return std::string("bm_upload: ") + names[t] + " is null";
    if (tables[t][0] != 0) return std::string("bm_upload: ") + names[t] + "[0] must be 0";
    for (int w = 0; w < n; w++)
      if (tables[t][w + 1] < tables[t][w])
        return std::string("bm_upload: ") + names[t] + " must not decrease";
  }
  if (h->awake_offset[n] > 0 && !h->awake_list) return "bm_upload: awake_list is null";
  if (h->manifold_offset[n] > 0 && (!h->a || !h->b || !h->pair || !h->sub || !h->row || !h->points))
    return "bm_upload: a manifold index column is null";
  for (int w = 0; w < n; w++) {
    const int nb = h->body_offset[w + 1] - h->body_offset[w];
    const int nm = h->manifold_offset[w + 1] - h->manifold_offset[w];
    const int np = h->pair_offset[w + 1] - h->pair_offset[w];
    const int moff = h->manifold_offset[w], aoff = h->awake_offset[w];
    if (h->substeps[w] < 1) return "bm_upload: substeps must be 1 or more";
    for (int i = aoff; i < h->awake_offset[w + 1]; i++)
      if (h->awake_list[i] < 0 || h->awake_list[i] >= nb) return "bm_upload: awake_list slot out of range";
    for (int m = moff; m < moff + nm; m++) {
      if (h->a[m] < 0 || h->a[m] >= nb || h->b[m] < 0 || h->b[m] >= nb)
        return "bm_upload: manifold body slot out of range";
      if (h->pair[m] < 0 || h->pair[m] >= np) return "bm_upload: manifold pair out of range";
      if (h->points[m] < 1 || h->points[m] > 4) return "bm_upload: a manifold has 1 to 4 points";
    }
  }
  return std::string();
}

// Whether starts, solved and row describe rows laid out by colour: runs in order within
// each world, solved at the resting run, row a permutation of each world's rows.
bool laid_out(const bm_columns* h) {
  if (!h->starts || !h->solved || !h->row) return false;
  std::vector<char> seen;
  for (int w = 0; w < h->n_worlds; w++) {
    const int nm = h->manifold_offset[w + 1] - h->manifold_offset[w], moff = h->manifold_offset[w];
    const int32_t* st = h->starts + BM_STARTS * w;
    if (st[0] != 0 || st[BM_STARTS - 1] != nm || st[BM_COLOURS + 1] != h->solved[w]) return false;
    for (int k = 0; k + 1 < BM_STARTS; k++)
      if (st[k + 1] < st[k]) return false;
    seen.assign(nm, 0);
    for (int m = moff; m < moff + nm; m++) {
      const int r = h->row[m];
      if (r < 0 || r >= nm || seen[r]) return false;
      seen[r] = 1;
    }
  }
  return true;
}

// Device memory of at least `bytes` at *p (capacity *cap), contents not kept.
int grow_device(bm_context* ctx, void** p, size_t* cap, size_t bytes, const char* what) {
  if (bytes <= *cap && *p) return BM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  size_t c = 65536;
  while (c < bytes) c *= 2;
  if (hipMalloc(p, c) != hipSuccess) return fail(ctx, BM_ERR_DEVICE, std::string(what) + ": hipMalloc");
  *cap = c;
  return BM_OK;
}

// Pinned host memory of at least `bytes` at *p (capacity *cap), contents not kept.
int grow_pinned(bm_context* ctx, void** p, size_t* cap, size_t bytes, const char* what) {
  if (bytes <= *cap && *p) return BM_OK;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  size_t c = 65536;
  while (c < bytes) c *= 2;
  if (hipHostMalloc(p, c, hipHostMallocDefault) != hipSuccess)
    return fail(ctx, BM_ERR_DEVICE, std::string(what) + ": hipHostMalloc");
  *cap = c;
  return BM_OK;
}

// The arena's column order: the arranged columns first, then the rest.
// The columns only the soft step itself writes and then reads, each one before it is read
// (prepare's levers, masses and Jacobians; the relax pass's separations): a context that
// keeps them on the device neither sends nor returns them.
const char* const DEVICE_ONLY[] = {"now", "ra", "rb", "normal_mass", "jac", "block", "lever", "anchor_a", "anchor_b",
                                   "fjac", "tangent_mass", "roll_mass1", "roll_mass2", "spin_mass"};

// The columns the step writes in every row before it reads them there (prepare: anchors,
// adjusted separations, first pushes and closing speeds; the solve: gap and sliding; rest()
// the resting rows' pushes, speeds and gap) and the host reads only after the step: a context
// that keeps the device-only columns on the device returns these but does not send them.
const char* const RESULTS_ONLY[] = {"la", "lb", "adjusted", "max_normal_impulse", "relative_velocity", "gap",
                                    "sliding"};

// The arena's column order: those arrange() moves first, then the rest, the results-only ones
// and last the device-only ones.
struct ArenaOrder {
  int order[NCOL];
  int arranged, results_only, device_only;
  ArenaOrder() {
    int k = 0;
    bool taken[NCOL] = {};
    for (const char* name : ARRANGED) {
      const int c = column_index(name);
      order[k++] = c;
      taken[c] = true;
    }
    arranged = k;
    for (const char* name : DEVICE_ONLY) taken[column_index(name)] = true;
    for (const char* name : RESULTS_ONLY) taken[column_index(name)] = true;
    for (int c = 0; c < NCOL; c++)
      if (!taken[c]) order[k++] = c;
    results_only = k;
    for (const char* name : RESULTS_ONLY) order[k++] = column_index(name);
    device_only = k;
    for (const char* name : DEVICE_ONLY) order[k++] = column_index(name);
  }
};

// The bodies' columns no step changes (src/world.js _write sets them as a body is added): with a
// body version set (bm_set_body_version) they live in a block of their own, sent only when the
// version changes.
const char* const CONSTANT[] = {"dims", "inertia", "inv_inertia", "gravity_scale", "damping", "angular_damping",
                                "type", "shape", "lock", "framed", "joint_head"};

bool is_constant(int c) {
  for (const char* name : CONSTANT)
    if (column_index(name) == c) return true;
  return false;
}

bool is_results_only(int c) {
  for (const char* name : RESULTS_ONLY)
    if (column_index(name) == c) return true;
  return false;
}

bool is_device_only(int c) {
  for (const char* name : DEVICE_ONLY)
    if (column_index(name) == c) return true;
  return false;
}

// Whether a shape record (kind first, a hull's number at 13) names a hull the table holds.
bool hull_ok(const bm_context* ctx, const double* r) {
  if (r[0] != 3) return true;
  const double h = r[13];
  return h >= 0 && h < ctx->hull_count && h == (double)(int32_t)h;
}

inline double host_min(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
inline double host_max(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

// src/hull.js:45-80 (the Hull constructor): each edge once, from the first face to reach
// it, with its two faces; the inscribed ball (inner) and the reach.
void derive_hull(int V, int F, const int32_t* start, const int32_t* loops, const double* v, const double* p,
                 std::vector<int32_t>& edges, std::vector<int32_t>& sides, double& inner, double& reach) {
  std::vector<int32_t> owner((size_t)V * V, -1);
  for (int f = 0; f < F; f++) {
    for (int i = start[f]; i < start[f + 1]; i++) {
      const int u = loops[i], w = loops[i + 1 < start[f + 1] ? i + 1 : start[f]];
      const int back = owner[(size_t)w * V + u];
      if (back < 0) owner[(size_t)u * V + w] = f;
      else { edges.push_back(w); edges.push_back(u); sides.push_back(back); sides.push_back(f); }
    }
  }
  reach = 0;
  for (int i = 0; i < V; i++) {
    const double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    reach = host_max(reach, std::sqrt(x * x + y * y + z * z));
  }
  inner = INFINITY;
  for (int f = 0; f < F; f++) {
    const double nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
    for (int i = start[f]; i < start[f + 1]; i++) {
      const int c = 3 * loops[i];
      inner = host_min(inner, nx * v[c] + ny * v[c + 1] + nz * v[c + 2]);
    }
  }
}

int grow(bm_context* ctx, Scratch& s, size_t bytes, const char* what) {
  if (bytes <= s.cap) return BM_OK;
  if (s.p) (void)hipFree(s.p);
  s.p = nullptr;
  size_t cap = 256;
  while (cap < bytes) cap *= 2;
  if (hipMalloc(&s.p, cap) != hipSuccess) return fail(ctx, BM_ERR_DEVICE, std::string(what) + ": hipMalloc");
  s.cap = cap;
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_create(bm_context** out, int32_t device) {
  if (!out) return BM_ERR_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return BM_ERR_DEVICE;
  bm_context* ctx = new bm_context();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess
      || hipEventCreate(&ctx->begin) != hipSuccess || hipEventCreate(&ctx->end) != hipSuccess
      || bm::level_noalloc_block(device, &ctx->noalloc_block) != hipSuccess) {
    bm_destroy(ctx);
    return BM_ERR_DEVICE;
  }
  *out = ctx;
  return BM_OK;
}

int bm_destroy(bm_context* ctx) {
  if (!ctx) return BM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->arena) (void)hipFree(ctx->arena);
  if (ctx->constants) (void)hipFree(ctx->constants);
  if (ctx->spare_arena) (void)hipFree(ctx->spare_arena);
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  if (ctx->pstage) (void)hipHostFree(ctx->pstage);
  for (Scratch* s : {&ctx->deg, &ctx->adj, &ctx->keys, &ctx->key_offset, &ctx->world_of, &ctx->pairs_in,
                     &ctx->pairs_out, &ctx->world_order})
    if (s->p) (void)hipFree(s->p);
  for (Scratch* s : {&ctx->hull_ints, &ctx->hull_doubles, &ctx->field_ints, &ctx->field_doubles, &ctx->level_scratch,
                     &ctx->colour_sort, &ctx->wide_bar, &ctx->wide_pick, &ctx->level_packed})
    if (s->p) (void)hipFree(s->p);
  bm_context::Broad& bb = ctx->broad;
  for (Scratch* s : {&bb.bodies, &bb.trees[0], &bb.trees[1], &bb.leaf, &bb.rows, &bb.out, &bb.hits, &bb.counters})
    if (s->p) (void)hipFree(s->p);
  bm::held_free(ctx->held);
  if (ctx->begin) (void)hipEventDestroy(ctx->begin);
  if (ctx->end) (void)hipEventDestroy(ctx->end);
  for (hipEvent_t e : ctx->piece)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return BM_OK;
}

const char* bm_last_error(const bm_context* ctx) { return ctx ? ctx->error.c_str() : "bm: no context"; }

int bm_held_load(bm_context* ctx, const bm_held* held) {
  if (!ctx) return BM_ERR_ARGUMENT;
  return bm::held_load(ctx->device, held, &ctx->held, ctx->error);
}

int bm_held_forward(bm_context* ctx, const double* state0, const double* mass, const double* friction,
                    const double* restitution, double* state_out) {
  if (!ctx || !state0 || !mass || !friction || !restitution || !state_out) return BM_ERR_ARGUMENT;
  return bm::held_forward(ctx->held, ctx->stream, state0, mass, friction, restitution, state_out, ctx->error);
}

int bm_held_backward(bm_context* ctx, const double* state_bar, double* state0_bar, double* mass_bar,
                     double* friction_bar, double* restitution_bar) {
  if (!ctx || !state_bar || !state0_bar || !mass_bar || !friction_bar || !restitution_bar) return BM_ERR_ARGUMENT;
  return bm::held_backward(ctx->held, ctx->stream, state_bar, state0_bar, mass_bar, friction_bar, restitution_bar,
                           ctx->error);
}

int bm_held_branches(bm_context* ctx, uint64_t* per_world) {
  if (!ctx || !per_world) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  return bm::held_branches(ctx->held, ctx->stream, per_world, ctx->error);
}

int bm_held_times(const bm_context* ctx, double* forward_ms, double* backward_ms) {
  if (!ctx) return BM_ERR_ARGUMENT;
  bm::held_times(ctx->held, forward_ms, backward_ms);
  return BM_OK;
}

double bm_last_kernel_ms(const bm_context* ctx) { return ctx ? ctx->last_ms : 0.0; }

int64_t bm_lds_limit(bm_context* ctx) {
  if (!ctx) return -1;
  if (ctx->lds_limit == 0) ctx->lds_limit = bm::soft_step_lds_limit(ctx->device);
  return (int64_t)ctx->lds_limit;
}

// Host copies between the caller's columns and the pinned staging: a large batch's (tens of
// MB and more) over up to 16 host threads, a small one's in turn. The list (a batch of
// thousands of worlds holds a copy per world and column, hundreds of thousands of small
// ones; a large copy is cut into 4 MB pieces) is cut into runs of about equal bytes, eight
// per thread, which the threads take in turn.
struct HostCopy {
  void* to;
  const void* from;
  size_t bytes;
};

// Host threads kept for the staging copies, made on first use and kept until the process ends
// (they wait on a condition variable between jobs), so a step's copies do not each start and
// join threads. One job at a time: a caller waits for the one before it to end.
class CopyThreads {
 public:
  explicit CopyThreads(int n) {
    for (int t = 0; t < n; t++) threads_.emplace_back([this] { loop(); });
  }
  int size() const { return (int)threads_.size(); }
  // The process's one set, never destroyed: its threads wait for work until the process ends.
  static CopyThreads& kept() {
    static CopyThreads* one =
        new CopyThreads(std::max(0, std::min(16, (int)std::thread::hardware_concurrency()) - 1));
    return *one;
  }
  // f on every kept thread and the caller's, returning when all have returned.
  void run(const std::function<void()>& f) {
    std::lock_guard<std::mutex> one(job_lock_);
    {
      std::lock_guard<std::mutex> l(m_);
      job_ = &f;
      busy_ = (int)threads_.size();
      gen_++;
    }
    start_.notify_all();
    f();
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void()>* f;
      {
        std::unique_lock<std::mutex> l(m_);
        start_.wait(l, [&] { return gen_ != seen; });
        seen = gen_;
        f = job_;
      }
      (*f)();
      std::lock_guard<std::mutex> l(m_);
      if (--busy_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex job_lock_, m_;
  std::condition_variable start_, done_;
  const std::function<void()>* job_ = nullptr;
  int busy_ = 0;
  uint64_t gen_ = 0;
};

// Uploads and downloads of at least this many bytes go in pieces (bm_context::PIECES).
constexpr size_t PIECED_BYTES = (size_t)8 << 20;

void run_copies(const std::vector<HostCopy>& copies) {
  size_t total = 0;
  for (const HostCopy& c : copies) total += c.bytes;
  const int threads = total >= ((size_t)4 << 20) ? std::max(1, std::min(16, (int)std::thread::hardware_concurrency())) : 1;
  if (threads <= 1) {
    for (const HostCopy& c : copies) std::memcpy(c.to, c.from, c.bytes);
    return;
  }
  constexpr size_t PIECE = (size_t)4 << 20;
  // Copies larger than a piece are cut into pieces; a list with none (a batch's many small
  // copies) is taken as it is.
  bool large = false;
  for (const HostCopy& c : copies) large = large || c.bytes > PIECE;
  std::vector<HostCopy> cut_up;
  if (large)
    for (const HostCopy& c : copies)
      for (size_t at = 0; at < c.bytes; at += PIECE)
        cut_up.push_back({static_cast<char*>(c.to) + at, static_cast<const char*>(c.from) + at, std::min(PIECE, c.bytes - at)});
  const std::vector<HostCopy>& pieces = large ? cut_up : copies;
  // Run r is pieces [cut[r], cut[r + 1]).
  const size_t runs = (size_t)threads * 8, goal = total / runs + 1;
  std::vector<size_t> cut(1, 0);
  size_t bytes = 0;
  for (size_t i = 0; i < pieces.size(); i++) {
    bytes += pieces[i].bytes;
    if (bytes >= goal * cut.size()) cut.push_back(i + 1);
  }
  if (cut.back() != pieces.size()) cut.push_back(pieces.size());
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t r; (r = next.fetch_add(1)) + 1 < cut.size();)
      for (size_t i = cut[r]; i < cut[r + 1]; i++) std::memcpy(pieces[i].to, pieces[i].from, pieces[i].bytes);
  };
  CopyThreads::kept().run(work);
}

// Units of kind u in one part (a bm_columns of one or more worlds).
size_t part_units(const bm_columns* h, Unit u) {
  const int n = h->n_worlds;
  switch (u) {
    case WORLD1: return (size_t)n + 1;
    case WORLD: return (size_t)n;
    case STARTS: return (size_t)n * BM_STARTS;
    case WORLD3: return (size_t)n * 3;
    case BODY: return (size_t)h->body_offset[n];
    case AWAKE: return (size_t)h->awake_offset[n];
    case MANIFOLD: return (size_t)h->manifold_offset[n];
    case KEPT: return (size_t)h->pair_offset[n];
  }
  return 0;
}

// The batch of every part's worlds in turn, into one arena: each column the parts' columns
// one after another (indices stay each world's own), the offset tables joined.
int upload_parts(bm_context* ctx, int nparts, const bm_columns* const* parts, const char* who) {
  ctx->uploaded = false;
  if (nparts <= 0) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": no worlds");
  int W = 0;
  bool laid = true;
  for (int q = 0; q < nparts; q++) {
    const bm_columns* h = parts[q];
    if (!h) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": null argument");
    for (int c = 0; c < NCOL; c++) {
      const Column& col = COLUMNS[c];
      if (col.unit <= WORLD3 && !field<const void*>(h, col.host))
        return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": " + col.name + " is null");
    }
    W += h->n_worlds;
  }
  // Each part's indices checked and its layout read, a batch's parts on the copy threads; the
  // first part found wanting (in order) names the error.
  std::vector<std::string> why(nparts);
  std::vector<uint8_t> part_laid(nparts, 0);
  auto check = [&](int q) {
    why[q] = check_part(parts[q]);
    part_laid[q] = why[q].empty() && laid_out(parts[q]) ? 1 : 0;
  };
  if (nparts >= 64) {
    std::atomic<int> next{0};
    CopyThreads::kept().run([&] {
      for (int q; (q = next.fetch_add(1)) < nparts;) check(q);
    });
  } else {
    for (int q = 0; q < nparts; q++) check(q);
  }
  for (int q = 0; q < nparts; q++) {
    if (!why[q].empty()) return fail(ctx, BM_ERR_ARGUMENT, why[q]);
    laid = laid && part_laid[q];
  }
  ctx->laid_out = laid;
  ctx->n_worlds = W;
  // The joined offset tables.
  std::vector<int32_t>* tables[4] = {&ctx->join_body, &ctx->join_manifold, &ctx->join_pair, &ctx->join_awake};
  for (auto* t : tables) t->assign(1, 0);
  for (int q = 0; q < nparts; q++) {
    const bm_columns* h = parts[q];
    const int32_t* src[4] = {h->body_offset, h->manifold_offset, h->pair_offset, h->awake_offset};
    for (int t = 0; t < 4; t++) {
      const int32_t base = tables[t]->back();
      for (int w = 1; w <= h->n_worlds; w++) tables[t]->push_back(base + src[t][w]);
    }
  }
  ctx->bodies = ctx->join_body.back();
  ctx->manifolds = ctx->join_manifold.back();
  ctx->pairs = ctx->join_pair.back();
  ctx->awake = ctx->join_awake.back();
  ctx->most_bodies = 0;
  for (int w = 0; w < W; w++) {
    const int32_t nb = ctx->join_body[w + 1] - ctx->join_body[w];
    if (nb > ctx->most_bodies) ctx->most_bodies = nb;
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), (std::string(who) + ": hipSetDevice").c_str());
  static const ArenaOrder ORDER;
  size_t at = 0, block = 0;
  const bool constants = ctx->body_version != 0;
  for (int k = 0; k < NCOL; k++) {
    const int c = ORDER.order[k];
    const Column& col = COLUMNS[c];
    const size_t bytes = units(ctx, col.unit) * col.per * col.elem;
    if (k == ORDER.device_only) ctx->device_only_at = at;
    if (k == ORDER.results_only) ctx->results_only_at = at;
    ctx->bytes[c] = bytes;
    ctx->in_block[c] = constants && is_constant(c);
    if (ctx->in_block[c]) {
      ctx->offset[c] = block;
      block += (bytes + 255) & ~(size_t)255;
    } else {
      ctx->offset[c] = at;
      at += (bytes + 255) & ~(size_t)255;
    }
    if (k + 1 == ORDER.arranged) ctx->arranged_end = at;
  }
  // The constant columns go up only when the block holds another version or other bodies.
  bool send_constants = constants && (ctx->kept_version != ctx->body_version || ctx->kept_bodies != ctx->bodies);
  ctx->kept_version = 0;  // set again below once the block holds this upload's constants
  if (constants && block > ctx->constants_cap) {
    if (ctx->constants) (void)hipFree(ctx->constants);
    ctx->constants = nullptr;
    ctx->constants_cap = 0;
    if (hipMalloc(&ctx->constants, block) != hipSuccess) return fail(ctx, BM_ERR_DEVICE, std::string(who) + ": hipMalloc");
    ctx->constants_cap = block;
    send_constants = true;
  }
  ctx->total = at;
  // What moves between host and device: all of it, or up to the device-only columns (and goes
  // up to the results-only ones).
  ctx->moved = ctx->keep_device_only ? ctx->device_only_at : ctx->total;
  const size_t sent = ctx->keep_device_only ? ctx->results_only_at : ctx->total;
  if (int rc = grow_device(ctx, &ctx->arena, &ctx->arena_cap, at, who)) return rc;
  if (int rc = grow_pinned(ctx, &ctx->stage, &ctx->stage_cap, at, who)) return rc;
  char* stage = static_cast<char*>(ctx->stage);
  const int32_t* joined[4] = {ctx->join_body.data(), ctx->join_manifold.data(), ctx->join_pair.data(),
                              ctx->join_awake.data()};
  std::vector<HostCopy> copies, constant_copies;
  if (send_constants)
    if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, block, who)) return rc;
  for (int c = 0; c < NCOL; c++) {
    const Column& col = COLUMNS[c];
    char* dst = (ctx->in_block[c] ? static_cast<char*>(ctx->pstage) : stage) + ctx->offset[c];
    if (ctx->in_block[c]) {
      for (int q = 0; q < nparts && send_constants; q++) {
        const size_t bytes = part_units(parts[q], col.unit) * col.per * col.elem;
        if (bytes == 0) continue;
        const void* from = field<const void*>(parts[q], col.host);
        if (!from) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": " + col.name + " is null");
        constant_copies.push_back({dst, from, bytes});
        dst += bytes;
      }
    } else if (col.unit == WORLD1) {
      const int t = c;  // the four offset tables are the first four columns
      std::memcpy(dst, joined[t], ctx->bytes[c]);
    } else if ((ctx->keep_device_only && (is_device_only(c) || is_results_only(c))) || col.unit == KEPT) {
      // left on the device (KEPT: written by the soft step before it is read)
    } else {
      for (int q = 0; q < nparts; q++) {
        const size_t bytes = part_units(parts[q], col.unit) * col.per * col.elem;
        if (bytes == 0) continue;
        const void* from = field<const void*>(parts[q], col.host);
        if (!from) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": " + col.name + " is null");
        copies.push_back({dst, from, bytes});
        dst += bytes;
      }
    }
    void* p = static_cast<char*>(ctx->in_block[c] ? ctx->constants : ctx->arena) + ctx->offset[c];
    ctx->buffer[c] = p;
    std::memcpy(reinterpret_cast<char*>(&ctx->dev) + col.dev, &p, sizeof(void*));
  }
  if (send_constants) {
    run_copies(constant_copies);
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->constants, ctx->pstage, block, hipMemcpyHostToDevice, ctx->stream),
                (std::string(who) + ": hipMemcpyAsync").c_str());
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), (std::string(who) + ": hipStreamSynchronize").c_str());
  }
  if (constants) {
    ctx->kept_version = ctx->body_version;
    ctx->kept_bodies = ctx->bodies;
  }
  {
    // The staging filled and sent in pieces cut at column starts: each piece goes up while the
    // next one is filled.
    constexpr int K = bm_context::PIECES;
    size_t cut[K + 1];
    cut[0] = 0;
    for (int k = 1; k <= K; k++) cut[k] = sent;
    for (int k = 1; k < K && sent >= PIECED_BYTES; k++) {
      const size_t goal = sent / K * k;
      size_t at = sent;
      for (int c = 0; c < NCOL; c++)
        if (!ctx->in_block[c] && ctx->offset[c] >= goal && ctx->offset[c] < at) at = ctx->offset[c];
      cut[k] = std::max(cut[k - 1], at);
    }
    std::vector<HostCopy> piece[K];
    for (const HostCopy& c : copies) {
      const size_t o = static_cast<size_t>(static_cast<char*>(c.to) - stage);
      int k = 0;
      while (k + 1 < K && o >= cut[k + 1]) k++;
      piece[k].push_back(c);
    }
    for (int k = 0; k < K; k++) {
      run_copies(piece[k]);
      if (cut[k + 1] > cut[k])
        HIP_OR_FAIL(ctx, hipMemcpyAsync(static_cast<char*>(ctx->arena) + cut[k], stage + cut[k], cut[k + 1] - cut[k],
                                        hipMemcpyHostToDevice, ctx->stream),
                    (std::string(who) + ": hipMemcpyAsync").c_str());
    }
  }
  ctx->moff_host = ctx->join_manifold;
  ctx->boff_host = ctx->join_body;
  ctx->dev.n_worlds = ctx->n_worlds;
  ctx->dev.world_order = nullptr;
  if (ctx->n_worlds > 1) {
    // The largest worlds first (longest-processing-time order): workgroups are handed out in
    // order as others finish, so an uneven batch's longest worlds do not start last.
    const int n = ctx->n_worlds;
    std::vector<int32_t>& order = ctx->order_host;
    order.resize(n);
    for (int w = 0; w < n; w++) order[w] = w;
    const std::vector<int32_t>& moff = ctx->moff_host;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return moff[x + 1] - moff[x] > moff[y + 1] - moff[y]; });
    if (int rc = grow(ctx, ctx->world_order, sizeof(int32_t) * (size_t)n, who)) return rc;
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->world_order.p, order.data(), sizeof(int32_t) * (size_t)n,
                                    hipMemcpyHostToDevice, ctx->stream),
                (std::string(who) + ": hipMemcpyAsync").c_str());
    ctx->dev.world_order = static_cast<const int*>(ctx->world_order.p);
  }
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), (std::string(who) + ": hipStreamSynchronize").c_str());
  ctx->uploaded = true;
  ctx->stored = false;
  return BM_OK;
}

// What the soft step stored by row (src/contact.js:1067-1088 store) into each part's pair
// columns at the row's pair: a solved row of a pair's own manifold (sub < 0) gives its pair's
// points, twist, friction and slip, and its points' ids and impulses; every other pair keeps
// what the part's arrays hold. Reads the part's manifold columns as the download left them.
void scatter_kept(const bm_context* ctx, int nparts, bm_columns* const* parts, const char* stage) {
  const auto col = [&](const char* name) { return stage + ctx->offset[column_index(name)]; };
  const uint8_t* kp = reinterpret_cast<const uint8_t*>(col("pair_points"));
  const uint32_t* kid = reinterpret_cast<const uint32_t*>(col("pair_pid"));
  const double* kimp = reinterpret_cast<const double*>(col("pair_normal_impulse"));
  const double* ktw = reinterpret_cast<const double*>(col("pair_twist"));
  const double* kfr = reinterpret_cast<const double*>(col("pair_friction"));
  const double* ksl = reinterpret_cast<const double*>(col("pair_slip"));
  size_t base = 0;  // the part's first row in the batch
  for (int q = 0; q < nparts; q++) {
    bm_columns* h = parts[q];
    for (int w = 0; w < h->n_worlds; w++) {
      const int m0 = h->manifold_offset[w], p0 = h->pair_offset[w], solved = h->solved[w];
      for (int lm = 0; lm < solved; lm++) {
        const int m = m0 + lm;
        if (h->sub[m] >= 0) continue;
        const size_t g = base + (size_t)m;
        const size_t p = (size_t)p0 + h->pair[m];
        const int count = kp[g] < 4 ? kp[g] : 4;
        h->pair_points[p] = kp[g];
        for (int k = 0; k < 3; k++) {
          h->pair_twist[3 * p + k] = ktw[3 * g + k];
          h->pair_friction[3 * p + k] = kfr[3 * g + k];
          h->pair_slip[3 * p + k] = ksl[3 * g + k];
        }
        for (int k = 0; k < count; k++) {
          h->pair_pid[4 * p + k] = kid[4 * g + k];
          h->pair_normal_impulse[4 * p + k] = kimp[4 * g + k];
        }
      }
    }
    base += (size_t)h->manifold_offset[h->n_worlds];
  }
}

// Every IO column back into the parts it came from.
int download_parts(bm_context* ctx, int nparts, bm_columns* const* parts, const char* who) {
  if (!ctx->uploaded) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": no batch uploaded");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), (std::string(who) + ": hipSetDevice").c_str());
  for (int q = 0; q < nparts; q++) {
    if (!parts[q]) return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": null argument");
    for (int c = 0; c < NCOL; c++)
      if (COLUMNS[c].dir == IO && part_units(parts[q], COLUMNS[c].unit) > 0 && !field<void*>(parts[q], COLUMNS[c].host) &&
          !(ctx->keep_device_only && is_device_only(c)))
        return fail(ctx, BM_ERR_ARGUMENT, std::string(who) + ": " + COLUMNS[c].name + " is null");
  }
  // The arena comes back in pieces cut at column starts, a column whole in one piece; each
  // piece's columns are scattered to the worlds while the pieces after it are in flight.
  constexpr int K = bm_context::PIECES;
  for (hipEvent_t& e : ctx->piece)
    if (!e) HIP_OR_FAIL(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), (std::string(who) + ": hipEventCreate").c_str());
  const size_t moved = ctx->moved;
  size_t cut[K + 1];
  cut[0] = 0;
  // A small world's arena comes back in one piece (each piece costs a copy, an event and a wait).
  for (int k = 1; k <= K; k++) cut[k] = moved;
  for (int k = 1; k < K && moved >= PIECED_BYTES; k++) {
    // the first column start at or after an equal share of the bytes
    const size_t goal = moved / K * k;
    size_t at = moved;
    for (int c = 0; c < NCOL; c++)
      if (!ctx->in_block[c] && ctx->offset[c] >= goal && ctx->offset[c] < at) at = ctx->offset[c];
    cut[k] = std::max(cut[k - 1], at);
  }
  char* stage = static_cast<char*>(ctx->stage);
  const char* arena = static_cast<const char*>(ctx->arena);
  for (int k = 0; k < K; k++) {
    if (cut[k + 1] == cut[k]) continue;
    HIP_OR_FAIL(ctx, hipMemcpyAsync(stage + cut[k], arena + cut[k], cut[k + 1] - cut[k], hipMemcpyDeviceToHost, ctx->stream),
                (std::string(who) + ": hipMemcpyAsync").c_str());
    HIP_OR_FAIL(ctx, hipEventRecord(ctx->piece[k], ctx->stream), (std::string(who) + ": hipEventRecord").c_str());
  }
  for (int k = 0; k < K; k++) {
    if (cut[k + 1] == cut[k]) continue;
    HIP_OR_FAIL(ctx, hipEventSynchronize(ctx->piece[k]), (std::string(who) + ": hipEventSynchronize").c_str());
    std::vector<HostCopy> copies;
    for (int c = 0; c < NCOL; c++) {
      if (COLUMNS[c].dir != IO || ctx->bytes[c] == 0 || COLUMNS[c].unit == KEPT) continue;
      if (ctx->keep_device_only && is_device_only(c)) continue;
      if (ctx->in_block[c] || ctx->offset[c] < cut[k] || ctx->offset[c] >= cut[k + 1]) continue;
      const char* from = stage + ctx->offset[c];
      for (int q = 0; q < nparts; q++) {
        const size_t bytes = part_units(parts[q], COLUMNS[c].unit) * COLUMNS[c].per * COLUMNS[c].elem;
        if (bytes == 0) continue;
        copies.push_back({field<void*>(parts[q], COLUMNS[c].host), from, bytes});
        from += bytes;
      }
    }
    run_copies(copies);
  }
  // Before a soft step ran on this upload the pairs keep what they held, as it left them.
  if (ctx->stored) scatter_kept(ctx, nparts, parts, stage);
  return BM_OK;
}

int bm_upload(bm_context* ctx, const bm_columns* h) {
  if (!ctx || !h) return fail(ctx, BM_ERR_ARGUMENT, "bm_upload: null argument");
  ctx->error.clear();
  return upload_parts(ctx, 1, &h, "bm_upload");
}

int bm_upload_worlds(bm_context* ctx, int32_t n, const bm_columns* const* worlds) {
  if (!ctx || !worlds) return fail(ctx, BM_ERR_ARGUMENT, "bm_upload_worlds: null argument");
  ctx->error.clear();
  return upload_parts(ctx, n, worlds, "bm_upload_worlds");
}

}  // extern "C"

namespace {

int colour_run(bm_context* ctx, bool capture, bool* captured);

// bm_soft_step's work; capture: the stream is capturing (bm_contact_phase), so the launch is
// queued and nothing waits (a batch of one world, which may take the cooperative grid-wide
// form, is never captured).
int soft_run(bm_context* ctx, bool capture) {
  if (!ctx) return BM_ERR_ARGUMENT;
  if (!ctx->uploaded) return fail(ctx, BM_ERR_ARGUMENT, "bm_soft_step: no batch uploaded");
  ctx->error.clear();
  if (!ctx->laid_out)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_soft_step: the batch's rows are not laid out by colour (call bm_colour)");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_soft_step: hipSetDevice");
  if (!capture) HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), "bm_soft_step: hipEventRecord");
  if (ctx->lds_limit == 0) ctx->lds_limit = bm::soft_step_lds_limit(ctx->device);
  // Velocities resident in LDS when the largest world's fit: 48 bytes a body.
  const size_t lds = (size_t)ctx->most_bodies * 6 * sizeof(double);
  (void)hipGetLastError();  // a stale error of an earlier call is not this launch's
  const int work = std::max(ctx->manifolds, ctx->awake);
  // A world whose velocities do not fit in LDS would sweep from HBM on one workgroup: the
  // whole grid pays off from a quarter of the threshold then (a landing pile's 4k-16k
  // manifolds: 28.5 -> 19.2 ms a step for its contact phase).
  const bool resident = ctx->resident && lds <= ctx->lds_limit;
  const int wide_at = resident ? ctx->wide_min : ctx->wide_min / 4;
  if (ctx->n_worlds == 1 && ctx->wide_min > 0 && ctx->manifolds >= wide_at) {
    // One world too large to sweep in one workgroup: the whole grid, cooperatively.
    if (!ctx->wide_bar.p) {
      if (int rc = grow(ctx, ctx->wide_bar, bm::WIDE_WORDS * sizeof(unsigned), "bm_soft_step")) return rc;
      HIP_OR_FAIL(ctx, hipMemset(ctx->wide_bar.p, 0, ctx->wide_bar.cap), "bm_soft_step: hipMemset");
    }
    if (int rc = grow(ctx, ctx->wide_pick, sizeof(int) * (size_t)std::max(ctx->manifolds, 1), "bm_soft_step")) return rc;
    const hipError_t e = bm::launch_soft_wide(ctx->dev, 0, work, static_cast<unsigned*>(ctx->wide_bar.p),
                                              static_cast<int*>(ctx->wide_pick.p), ctx->device, ctx->stream,
                                              &ctx->last_wide);
    if (e != hipSuccess) {
      // A cooperative launch can be refused while another process holds the device's
      // cooperative resources: the step then takes the one-workgroup path, with the same bits.
      (void)hipGetLastError();
      if (ctx->wide_refused++ == 0)
        std::fprintf(stderr, "bearing_amd: the grid-wide soft step was refused (%s); this context steps its worlds "
                     "one workgroup each while it is\n", hipGetErrorString(e));
      ctx->wide_error = hipGetErrorString(e);
      ctx->last_wide = 0;
      bm::launch_soft_step(ctx->dev, ctx->n_worlds, ctx->shape, resident ? lds : 0, ctx->stream);
    }
  } else {
    ctx->last_wide = 0;
    bm::launch_soft_step(ctx->dev, ctx->n_worlds, ctx->shape, resident ? lds : 0, ctx->stream);
  }
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_soft_step: launch");
  ctx->stored = true;
  if (capture) return BM_OK;
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), "bm_soft_step: hipEventRecord");
  HIP_OR_FAIL(ctx, hipEventSynchronize(ctx->end), "bm_soft_step: kernel");
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_soft_step: hipEventElapsedTime");
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_soft_step(bm_context* ctx) { return soft_run(ctx, false); }

int bm_set_graph(bm_context* ctx, int32_t on) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  ctx->graph_on = on != 0;
  return BM_OK;
}

int bm_graph_counts(const bm_context* ctx, int64_t* out) {
  if (!ctx || !out) return BM_ERR_ARGUMENT;
  out[0] = ctx->graph_launches;
  out[1] = ctx->graph_instantiations;
  out[2] = ctx->graph_updates;
  return BM_OK;
}

int bm_contact_phase(bm_context* ctx) {
  if (!ctx) return BM_ERR_ARGUMENT;
  if (!ctx->graph_on || ctx->n_worlds < 2) {
    if (int rc = colour_run(ctx, false, nullptr)) return rc;
    return soft_run(ctx, false);
  }
  // The colouring's and the soft step's launches captured as a graph: the executable graph of
  // the step before is updated to this step's arguments when the shapes allow, else made anew.
  bool captured = false;
  int rc = colour_run(ctx, true, &captured);
  if (rc == BM_OK && captured) rc = soft_run(ctx, true);
  if (!captured) {
    if (rc != BM_OK) return rc;
    return soft_run(ctx, false);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ended = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != BM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ended != hipSuccess) return fail(ctx, BM_ERR_DEVICE, std::string("bm_contact_phase: hipStreamEndCapture: ") + hipGetErrorString(ended));
  bool ready = false;
  if (ctx->graph_exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult why;
    if (hipGraphExecUpdate(ctx->graph_exec, graph, &bad, &why) == hipSuccess) {
      ready = true;
      ctx->graph_updates++;
    } else {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(ctx->graph_exec);
      ctx->graph_exec = nullptr;
    }
  }
  if (!ready) {
    const hipError_t e = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      ctx->graph_exec = nullptr;
      (void)hipGraphDestroy(graph);
      return fail(ctx, BM_ERR_DEVICE, std::string("bm_contact_phase: hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    ctx->graph_instantiations++;
  }
  (void)hipGraphDestroy(graph);
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), "bm_contact_phase: hipEventRecord");
  HIP_OR_FAIL(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream), "bm_contact_phase: hipGraphLaunch");
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), "bm_contact_phase: hipEventRecord");
  ctx->graph_launches++;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_contact_phase: kernel");
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_contact_phase: hipEventElapsedTime");
  return BM_OK;
}

int bm_download(bm_context* ctx, bm_columns* h) {
  if (!ctx || !h) return fail(ctx, BM_ERR_ARGUMENT, "bm_download: null argument");
  ctx->error.clear();
  return download_parts(ctx, 1, &h, "bm_download");
}

int bm_download_worlds(bm_context* ctx, int32_t n, bm_columns* const* worlds) {
  if (!ctx || !worlds) return fail(ctx, BM_ERR_ARGUMENT, "bm_download_worlds: null argument");
  ctx->error.clear();
  if (n <= 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_download_worlds: no worlds");
  return download_parts(ctx, n, worlds, "bm_download_worlds");
}

}  // extern "C"

namespace {

// bm_colour's work. capture: once the host's part is done and its one copy queued, the stream
// starts capturing (for bm_contact_phase's graph), and the launches are queued without a wait;
// *captured says whether it did (not for the streamed path, whose sort allocates as it goes).
int colour_run(bm_context* ctx, bool capture, bool* captured) {
  if (captured) *captured = false;
  if (!ctx) return BM_ERR_ARGUMENT;
  if (!ctx->uploaded) return fail(ctx, BM_ERR_ARGUMENT, "bm_colour: no batch uploaded");
  ctx->error.clear();
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_colour: hipSetDevice");
  const int n = ctx->n_worlds;
  // Each world's sort width: a power of two at least its manifold count.
  const std::vector<int32_t>& moff = ctx->moff_host;
  ctx->key_offsets.assign(n + 1, 0);
  for (int w = 0; w < n; w++) {
    int width = 1;
    while (width < moff[w + 1] - moff[w]) width *= 2;
    ctx->key_offsets[w + 1] = ctx->key_offsets[w] + width;
  }
  const size_t manifolds = ctx->manifolds;
  if (int rc = grow(ctx, ctx->deg, ((size_t)ctx->bodies + n) * sizeof(int32_t), "bm_colour")) return rc;
  if (int rc = grow(ctx, ctx->adj, 2 * manifolds * sizeof(int32_t) + 4, "bm_colour")) return rc;
  if (int rc = grow(ctx, ctx->keys, (size_t)ctx->key_offsets[n] * sizeof(uint64_t), "bm_colour")) return rc;
  if (int rc = grow(ctx, ctx->key_offset, (n + 1) * sizeof(int32_t), "bm_colour")) return rc;
  if (int rc = grow(ctx, ctx->world_of, manifolds * sizeof(int32_t) + 4, "bm_colour")) return rc;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->key_offset.p, ctx->key_offsets.data(), (n + 1) * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream), "bm_colour: hipMemcpyAsync");
  auto buf = [&](const char* name) { return ctx->buffer[column_index(name)]; };
  bm::ColourDev cd{};
  cd.n_worlds = n;
  cd.body_offset = static_cast<const int*>(buf("body_offset"));
  cd.manifold_offset = static_cast<const int*>(buf("manifold_offset"));
  cd.a = static_cast<const int*>(buf("a"));
  cd.b = static_cast<const int*>(buf("b"));
  cd.type = static_cast<const uint8_t*>(buf("type"));
  cd.frozen = static_cast<const uint8_t*>(buf("frozen"));
  cd.colour = static_cast<int*>(buf("colour"));
  cd.order = static_cast<int*>(buf("order"));
  cd.row = static_cast<int*>(buf("row"));
  cd.starts = static_cast<int*>(buf("starts"));
  cd.solved = static_cast<int*>(buf("solved"));
  cd.deg = static_cast<int*>(ctx->deg.p);
  cd.adj = static_cast<int*>(ctx->adj.p);
  cd.keys = static_cast<uint64_t*>(ctx->keys.p);
  cd.key_offset = static_cast<const int*>(ctx->key_offset.p);
  cd.world_order = ctx->dev.world_order;
  // Colour in LDS when every world's data fits there.
  if (ctx->colour_lds_limit == 0) ctx->colour_lds_limit = bm::colour_lds_limit(ctx->device);
  size_t lds = 0;
  for (int w = 0; w < n; w++) {
    const size_t need = bm::colour_lds_bytes(ctx->boff_host[w + 1] - ctx->boff_host[w], moff[w + 1] - moff[w],
                                             ctx->key_offsets[w + 1] - ctx->key_offsets[w]);
    if (need > lds) lds = need;
  }
  cd.in_lds = ctx->colour_in_lds == 1 && lds <= ctx->colour_lds_limit ? 1 : 0;
  (void)hipGetLastError();
  const int bodies0 = ctx->boff_host[1] - ctx->boff_host[0];
  const bool streamed = !cd.in_lds && n == 1 && ctx->colour_in_lds != 0 && manifolds > 0 &&
                        bm::colour_stream_lds(bodies0) <= ctx->colour_lds_limit;
  if (int rc = grow_device(ctx, &ctx->spare_arena, &ctx->spare_cap, ctx->arranged_end, "bm_colour")) return rc;
  if (capture && !streamed) {
    HIP_OR_FAIL(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                "bm_contact_phase: hipStreamBeginCapture");
    *captured = true;
  }
  bm::launch_world_of(cd.manifold_offset, n, static_cast<int*>(ctx->world_of.p), ctx->stream);
  if (streamed) {
    // One world too large to colour in LDS whose bodies' masks fit: streamed, then sorted.
    const size_t temp = bm::colour_stream_scratch((int)manifolds);
    const size_t need = manifolds * (8 + 8 + 4 + 4) + 16 + temp + 256;
    if (int rc = grow(ctx, ctx->colour_sort, need, "bm_colour")) return rc;
    if (bm::launch_colour_stream(cd, 0, (int)manifolds, bodies0, ctx->colour_sort.p, need, ctx->stream) != 0)
      return fail(ctx, BM_ERR_DEVICE, "bm_colour: the radix sort could not run");
    ctx->last_colour_path = 2;
  } else {
    bm::launch_colour(cd, lds, ctx->stream);
    ctx->last_colour_path = cd.in_lds ? 1 : 0;
  }
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_colour: launch");
  // Arrange: every column narrow() writes, into the spare arena at the same offsets, then
  // the arranged block is copied back in one piece.
  bm::ArrangeDev ad{};
  ad.manifolds = (long)manifolds;
  ad.manifold_offset = cd.manifold_offset;
  ad.order = cd.order;
  ad.world_of = static_cast<const int*>(ctx->world_of.p);
  for (int k = 0; k < NARRANGED; k++) {
    const int c = column_index(ARRANGED[k]);
    ad.columns[k] = bm::ArrangeColumn{ctx->buffer[c], static_cast<char*>(ctx->spare_arena) + ctx->offset[c],
                                      (int)COLUMNS[c].elem, COLUMNS[c].per};
  }
  bm::launch_arrange(ad, NARRANGED, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_colour: arrange");
  if (manifolds > 0)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->arena, ctx->spare_arena, ctx->arranged_end, hipMemcpyDeviceToDevice,
                                    ctx->stream), "bm_colour: hipMemcpyAsync");
  bm::RestDev rd{};
  rd.manifolds = (long)manifolds;
  rd.manifold_offset = cd.manifold_offset;
  rd.solved = cd.solved;
  rd.world_of = ad.world_of;
  rd.points = static_cast<const uint8_t*>(buf("points"));
  rd.separation = static_cast<const double*>(buf("separation"));
  rd.normal_impulse = static_cast<const double*>(buf("normal_impulse"));
  rd.max_normal_impulse = static_cast<double*>(buf("max_normal_impulse"));
  rd.relative_velocity = static_cast<double*>(buf("relative_velocity"));
  rd.gap = static_cast<double*>(buf("gap"));
  bm::launch_rest(rd, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_colour: rest");
  if (!capture) HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_colour: kernel");
  ctx->laid_out = true;
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_colour(bm_context* ctx) { return colour_run(ctx, false, nullptr); }

int bm_collide_host(bm_context* ctx, int32_t n, const double* a, const double* b, const double* margin,
                    const double* tolerance, int32_t* count, double* normal, double* points, double* separations,
                    uint32_t* ids, double* loss) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_collide_host: n must be 0 or more");
  if (n == 0) return BM_OK;
  if (!a || !b || !margin || !tolerance || !count || !normal || !points || !separations || !ids)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_collide_host: null argument");
  for (int32_t k = 0; k < n; k++) {
    const double ka = a[19 * (size_t)k], kb = b[19 * (size_t)k];
    if (!(ka == 0 || ka == 1 || ka == 2 || ka == 3) || !(kb == 0 || kb == 1 || kb == 2 || kb == 3))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_collide_host: a shape's kind must be 0, 1, 2 or 3");
    if (!hull_ok(ctx, a + 19 * (size_t)k) || !hull_ok(ctx, b + 19 * (size_t)k))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_collide_host: a hull's number is not in the hull table (bm_set_hulls)");
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_collide_host: hipSetDevice");
  const size_t shapes = 19 * sizeof(double) * (size_t)n, scalars = sizeof(double) * (size_t)n;
  const size_t in_bytes = 2 * shapes + 2 * scalars;
  const size_t out_bytes = 184 * (size_t)n + 64;
  if (int rc = grow(ctx, ctx->pairs_in, in_bytes, "bm_collide_host")) return rc;
  if (int rc = grow(ctx, ctx->pairs_out, out_bytes, "bm_collide_host")) return rc;
  char* in = static_cast<char*>(ctx->pairs_in.p);
  char* out = static_cast<char*>(ctx->pairs_out.p);
  bm::CollideDev d{};
  d.n = n;
  d.a = reinterpret_cast<const double*>(in);
  d.b = reinterpret_cast<const double*>(in + shapes);
  d.margin = reinterpret_cast<const double*>(in + 2 * shapes);
  d.tolerance = reinterpret_cast<const double*>(in + 2 * shapes + scalars);
  d.normal = reinterpret_cast<double*>(out);
  d.points = reinterpret_cast<double*>(out + 3 * 8 * (size_t)n);
  d.separations = reinterpret_cast<double*>(out + 15 * 8 * (size_t)n);
  d.count = reinterpret_cast<int32_t*>(out + 19 * 8 * (size_t)n);
  d.ids = reinterpret_cast<uint32_t*>(out + 19 * 8 * (size_t)n + 4 * (size_t)n);
  d.loss = reinterpret_cast<double*>(out + 176 * (size_t)n);
  d.hulls = ctx->hulls;
  // One copy each way through pinned staging: inputs as laid out on the device, then the outputs.
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, in_bytes + out_bytes, "bm_collide_host")) return rc;
  char* st = static_cast<char*>(ctx->pstage);
  std::memcpy(st, a, shapes);
  std::memcpy(st + shapes, b, shapes);
  std::memcpy(st + 2 * shapes, margin, scalars);
  std::memcpy(st + 2 * shapes + scalars, tolerance, scalars);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(in, st, in_bytes, hipMemcpyHostToDevice, ctx->stream), "bm_collide_host: copy");
  (void)hipGetLastError();
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), "bm_collide_host: hipEventRecord");
  bm::launch_collide(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_collide_host: launch");
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), "bm_collide_host: hipEventRecord");
  const size_t produced = 184 * (size_t)n;
  char* back = st + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, out, produced, hipMemcpyDeviceToHost, ctx->stream), "bm_collide_host: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_collide_host: kernel");
  std::memcpy(normal, back, 24 * (size_t)n);
  std::memcpy(points, back + 3 * 8 * (size_t)n, 96 * (size_t)n);
  std::memcpy(separations, back + 15 * 8 * (size_t)n, 32 * (size_t)n);
  std::memcpy(count, back + 19 * 8 * (size_t)n, 4 * (size_t)n);
  if (loss) std::memcpy(loss, back + 176 * (size_t)n, 8 * (size_t)n);
  std::memcpy(ids, back + 19 * 8 * (size_t)n + 4 * (size_t)n, 16 * (size_t)n);
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_collide_host: hipEventElapsedTime");
  return BM_OK;
}

int bm_meet_host(bm_context* ctx, int32_t n, const double* a, const double* b, const double* pair, int32_t* count,
                 double* normal, double* points, double* separations, uint32_t* ids, double* shift, int32_t* flags,
                 double* loss) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_host: n must be 0 or more");
  if (n == 0) return BM_OK;
  if (!a || !b || !pair || !count || !normal || !points || !separations || !ids || !shift || !flags)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_host: null argument");
  for (int32_t k = 0; k < n; k++) {
    const double ka = a[32 * (size_t)k], kb = b[32 * (size_t)k];
    if (!(ka == 0 || ka == 1 || ka == 2 || ka == 3) || !(kb == 0 || kb == 1 || kb == 2 || kb == 3))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_host: a shape's kind must be 0, 1, 2 or 3");
    if (!hull_ok(ctx, a + 32 * (size_t)k) || !hull_ok(ctx, b + 32 * (size_t)k))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_host: a hull's number is not in the hull table (bm_set_hulls)");
    if (!(pair[6 * (size_t)k + 2] > 0)) return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_host: dt must be positive");
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_meet_host: hipSetDevice");
  const size_t bodies = 32 * sizeof(double) * (size_t)n, pairs = 6 * sizeof(double) * (size_t)n;
  const size_t in_bytes = 2 * bodies + pairs;
  const size_t out_bytes = 280 * (size_t)n + 64;
  if (int rc = grow(ctx, ctx->pairs_in, in_bytes, "bm_meet_host")) return rc;
  if (int rc = grow(ctx, ctx->pairs_out, out_bytes, "bm_meet_host")) return rc;
  char* in = static_cast<char*>(ctx->pairs_in.p);
  char* out = static_cast<char*>(ctx->pairs_out.p);
  const size_t N = (size_t)n;
  bm::MeetDev d{};
  d.n = n;
  d.a = reinterpret_cast<const double*>(in);
  d.b = reinterpret_cast<const double*>(in + bodies);
  d.pair = reinterpret_cast<const double*>(in + 2 * bodies);
  d.normal = reinterpret_cast<double*>(out);
  d.points = reinterpret_cast<double*>(out + 24 * N);
  d.separations = reinterpret_cast<double*>(out + 120 * N);
  d.shift = reinterpret_cast<double*>(out + 152 * N);
  d.count = reinterpret_cast<int32_t*>(out + 248 * N);
  d.flags = reinterpret_cast<int32_t*>(out + 252 * N);
  d.ids = reinterpret_cast<uint32_t*>(out + 256 * N);
  d.loss = reinterpret_cast<double*>(out + 272 * N);
  d.hulls = ctx->hulls;
  // One copy each way through pinned staging.
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, in_bytes + out_bytes, "bm_meet_host")) return rc;
  char* st = static_cast<char*>(ctx->pstage);
  run_copies({{st, a, bodies}, {st + bodies, b, bodies}, {st + 2 * bodies, pair, pairs}});
  HIP_OR_FAIL(ctx, hipMemcpyAsync(in, st, in_bytes, hipMemcpyHostToDevice, ctx->stream), "bm_meet_host: copy");
  (void)hipGetLastError();
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), "bm_meet_host: hipEventRecord");
  bm::launch_meet(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_meet_host: launch");
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), "bm_meet_host: hipEventRecord");
  char* back = st + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, out, 280 * N, hipMemcpyDeviceToHost, ctx->stream), "bm_meet_host: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_meet_host: kernel");
  std::vector<HostCopy> outs = {{normal, back, 24 * N}, {points, back + 24 * N, 96 * N},
                                {separations, back + 120 * N, 32 * N}, {shift, back + 152 * N, 96 * N},
                                {count, back + 248 * N, 4 * N}, {flags, back + 252 * N, 4 * N},
                                {ids, back + 256 * N, 16 * N}};
  if (loss) outs.push_back({loss, back + 272 * N, 8 * N});
  run_copies(outs);
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_meet_host: hipEventElapsedTime");
  return BM_OK;
}

}  // extern "C"

namespace {

// The common start of bm_meet_indexed and bm_meet_packed: the arguments checked, pairs_out
// grown to out_bytes (the pass's outputs first, 280 bytes a pair), the records and the table
// sent, the pass launched (timed by ctx's events); the staging for the results starts at
// *in_bytes_out in ctx->pstage.
int meet_start(bm_context* ctx, const char* who, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
               const int32_t* ib, const double* pair, size_t out_bytes, bm::MeetDev& d, size_t* in_bytes_out) {
  const std::string w(who);
  if (!table || !ia || !ib || !pair) return fail(ctx, BM_ERR_ARGUMENT, w + ": null argument");
  // The records and pairs checked in chunks (a large pass's on the copy threads); the first
  // chunk found wanting names the error.
  const int64_t items = (int64_t)bodies + n, chunk = 16384, chunks = (items + chunk - 1) / chunk;
  std::vector<uint8_t> why((size_t)chunks, 0);  // 0 good; 1 kind, 2 hull, 3 body, 4 dt
  auto check = [&](int64_t c) {
    for (int64_t i = c * chunk, e = std::min(items, (c + 1) * chunk); i < e && !why[c]; i++) {
      if (i < bodies) {
        const double kind = table[32 * (size_t)i];
        if (!(kind == 0 || kind == 1 || kind == 2 || kind == 3)) why[c] = 1;
        else if (!hull_ok(ctx, table + 32 * (size_t)i)) why[c] = 2;
      } else {
        const int64_t k = i - bodies;
        if (ia[k] < 0 || ia[k] >= bodies || ib[k] < 0 || ib[k] >= bodies) why[c] = 3;
        else if (!(pair[6 * (size_t)k + 2] > 0)) why[c] = 4;
      }
    }
  };
  if (chunks >= 4) {
    std::atomic<int64_t> next{0};
    CopyThreads::kept().run([&] {
      for (int64_t c; (c = next.fetch_add(1)) < chunks;) check(c);
    });
  } else {
    for (int64_t c = 0; c < chunks; c++) check(c);
  }
  static const char* const WHY[] = {"", ": a shape's kind must be 0, 1, 2 or 3",
                                    ": a hull's number is not in the hull table (bm_set_hulls)",
                                    ": a pair names a body the table lacks", ": dt must be positive"};
  for (int64_t c = 0; c < chunks; c++)
    if (why[c]) return fail(ctx, BM_ERR_ARGUMENT, w + WHY[why[c]]);
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), (w + ": hipSetDevice").c_str());
  const size_t N = (size_t)n, T = 32 * sizeof(double) * (size_t)bodies;
  const size_t pairs = 6 * sizeof(double) * N, idx = (4 * N + 255) & ~(size_t)255;
  const size_t in_bytes = T + pairs + 2 * idx;
  if (int rc = grow(ctx, ctx->pairs_in, in_bytes, who)) return rc;
  if (int rc = grow(ctx, ctx->pairs_out, out_bytes, who)) return rc;
  char* in = static_cast<char*>(ctx->pairs_in.p);
  char* out = static_cast<char*>(ctx->pairs_out.p);
  d.n = n;
  d.a = reinterpret_cast<const double*>(in);
  d.b = d.a;
  d.pair = reinterpret_cast<const double*>(in + T);
  d.ia = reinterpret_cast<const int*>(in + T + pairs);
  d.ib = reinterpret_cast<const int*>(in + T + pairs + idx);
  d.normal = reinterpret_cast<double*>(out);
  d.points = reinterpret_cast<double*>(out + 24 * N);
  d.separations = reinterpret_cast<double*>(out + 120 * N);
  d.shift = reinterpret_cast<double*>(out + 152 * N);
  d.count = reinterpret_cast<int32_t*>(out + 248 * N);
  d.flags = reinterpret_cast<int32_t*>(out + 252 * N);
  d.ids = reinterpret_cast<uint32_t*>(out + 256 * N);
  d.loss = reinterpret_cast<double*>(out + 272 * N);
  d.hulls = ctx->hulls;
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, in_bytes + 400 * N + 4096, who)) return rc;
  char* st = static_cast<char*>(ctx->pstage);
  // Staged and sent in pieces when large (a batch's table of every meeting body): each piece
  // goes up while the next one is staged.
  const std::vector<HostCopy> all = {
      {st, table, T}, {st + T, pair, pairs}, {st + T + pairs, ia, 4 * N}, {st + T + pairs + idx, ib, 4 * N}};
  constexpr int K = bm_context::PIECES;
  const int pieces = in_bytes >= PIECED_BYTES ? K : 1;
  for (int k = 0; k < pieces; k++) {
    const size_t lo = in_bytes / pieces * k, hi = k + 1 == pieces ? in_bytes : in_bytes / pieces * (k + 1);
    std::vector<HostCopy> part;
    for (const HostCopy& c : all) {
      const size_t at = static_cast<size_t>(static_cast<char*>(c.to) - st);
      const size_t a = std::max(at, lo), b = std::min(at + c.bytes, hi);
      if (a < b) part.push_back({st + a, static_cast<const char*>(c.from) + (a - at), b - a});
    }
    run_copies(part);
    HIP_OR_FAIL(ctx, hipMemcpyAsync(in + lo, st + lo, hi - lo, hipMemcpyHostToDevice, ctx->stream), (w + ": copy").c_str());
  }
  (void)hipGetLastError();
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), (w + ": hipEventRecord").c_str());
  bm::launch_meet(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), (w + ": launch").c_str());
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), (w + ": hipEventRecord").c_str());
  *in_bytes_out = in_bytes;
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_meet_indexed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                    const int32_t* ib, const double* pair, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids, double* shift, int32_t* flags, double* loss) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || bodies < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_indexed: n and bodies must be 0 or more");
  if (n == 0) return BM_OK;
  if (!count || !normal || !points || !separations || !ids || !shift || !flags)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_indexed: null argument");
  bm::MeetDev d{};
  size_t in_bytes = 0;
  const size_t N = (size_t)n;
  if (int rc = meet_start(ctx, "bm_meet_indexed", bodies, table, n, ia, ib, pair, 280 * N + 64, d, &in_bytes))
    return rc;
  char* st = static_cast<char*>(ctx->pstage);
  char* back = st + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, ctx->pairs_out.p, 280 * N, hipMemcpyDeviceToHost, ctx->stream),
              "bm_meet_indexed: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_meet_indexed: kernel");
  std::vector<HostCopy> outs = {{normal, back, 24 * N}, {points, back + 24 * N, 96 * N},
                                {separations, back + 120 * N, 32 * N}, {shift, back + 152 * N, 96 * N},
                                {count, back + 248 * N, 4 * N}, {flags, back + 252 * N, 4 * N},
                                {ids, back + 256 * N, 16 * N}};
  if (loss) outs.push_back({loss, back + 272 * N, 8 * N});
  run_copies(outs);
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_meet_indexed: hipEventElapsedTime");
  return BM_OK;
}

int bm_meet_packed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                   const int32_t* ib, const double* pair, int32_t* count, int32_t* flags, int32_t* row,
                   double* normal, double* points, double* separations, uint32_t* ids, double* shift, double* loss) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || bodies < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_packed: n and bodies must be 0 or more");
  if (n == 0) return BM_OK;
  if (!count || !flags || !row || !normal || !points || !separations || !ids || !shift || !loss)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_meet_packed: null argument");
  bm::MeetDev d{};
  size_t in_bytes = 0;
  const size_t N = (size_t)n;
  // The packed columns after the pass's own outputs, in the same scratch: the rows (N + 1),
  // then room for N found rows of each column.
  const size_t pk_row = 280 * N + 64, pk_normal = pk_row + ((4 * (N + 1) + 255) & ~(size_t)255);
  const size_t pk_points = pk_normal + 24 * N, pk_seps = pk_points + 96 * N, pk_shift = pk_seps + 32 * N;
  const size_t pk_loss = pk_shift + 96 * N, pk_ids = pk_loss + 8 * N, pk_sums = pk_ids + 16 * N;
  const size_t pk_end = pk_sums + 4 * (N / 256 + 2);
  if (int rc = meet_start(ctx, "bm_meet_packed", bodies, table, n, ia, ib, pair, pk_end, d, &in_bytes)) return rc;
  char* out = static_cast<char*>(ctx->pairs_out.p);
  bm::MeetPacked p{};
  p.row = reinterpret_cast<int*>(out + pk_row);
  p.found = p.row + N;
  p.normal = reinterpret_cast<double*>(out + pk_normal);
  p.points = reinterpret_cast<double*>(out + pk_points);
  p.separations = reinterpret_cast<double*>(out + pk_seps);
  p.shift = reinterpret_cast<double*>(out + pk_shift);
  p.loss = reinterpret_cast<double*>(out + pk_loss);
  p.ids = reinterpret_cast<uint32_t*>(out + pk_ids);
  p.sums = reinterpret_cast<int*>(out + pk_sums);
  bm::launch_meet_pack(d, p, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_meet_packed: pack");
  // Every pair's count, flags and row first, then only the rows found.
  char* back = static_cast<char*>(ctx->pstage) + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, d.count, 8 * N, hipMemcpyDeviceToHost, ctx->stream), "bm_meet_packed: copy");
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back + 8 * N, p.row, 4 * (N + 1), hipMemcpyDeviceToHost, ctx->stream),
              "bm_meet_packed: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_meet_packed: kernel");
  const int32_t* bi = reinterpret_cast<const int32_t*>(back);
  const size_t found = (size_t)bi[2 * N + N];
  if (found > 0) {
    const struct {
      void* to;
      const void* from;
      size_t bytes;
    } cols[] = {{normal, p.normal, 24 * found}, {points, p.points, 96 * found}, {separations, p.separations, 32 * found},
                {shift, p.shift, 96 * found},   {loss, p.loss, 8 * found},       {ids, p.ids, 16 * found}};
    size_t at = (12 * N + 4 + 255) & ~(size_t)255;
    for (const auto& c : cols) {
      HIP_OR_FAIL(ctx, hipMemcpyAsync(back + at, c.from, c.bytes, hipMemcpyDeviceToHost, ctx->stream),
                  "bm_meet_packed: copy");
      at += (c.bytes + 255) & ~(size_t)255;
    }
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_meet_packed: copy");
    std::vector<HostCopy> copies;
    at = (12 * N + 4 + 255) & ~(size_t)255;
    for (const auto& c : cols) {
      copies.push_back({c.to, back + at, c.bytes});
      at += (c.bytes + 255) & ~(size_t)255;
    }
    run_copies(copies);
  }
  std::vector<HostCopy> heads = {{count, back, 4 * N}, {flags, back + 4 * N, 4 * N}, {row, back + 8 * N, 4 * N}};
  run_copies(heads);
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_meet_packed: hipEventElapsedTime");
  return BM_OK;
}

int bm_set_hulls(bm_context* ctx, int32_t n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l,
                 const double* vertices, const int32_t* start, const int32_t* loops, const double* planes) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || (n > 0 && (!hull_v || !hull_f || !hull_l || !vertices || !start || !loops || !planes)))
    return fail(ctx, BM_ERR_ARGUMENT, "bm_set_hulls: null argument");
  // The reference's unit cube (src/collide.js:1298, as src/hull.js build() lays it out), last.
  static const double CUBE_V[24] = {-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5,
                                    -0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5};
  static const int32_t CUBE_S[6] = {0, 4, 8, 12, 16, 20};
  static const int32_t CUBE_L[24] = {0, 1, 5, 4, 0, 4, 6, 2, 0, 2, 3, 1, 1, 3, 7, 5, 4, 5, 7, 6, 2, 6, 7, 3};
  static const double CUBE_P[24] = {0, -1, 0, 0.5, -1, 0, 0, 0.5, 0, 0, -1, 0.5, 1, 0, 0, 0.5, 0, 0, 1, 0.5, 0, 1, 0, 0.5};
  std::vector<int32_t> info, ints_start, ints_loops, ints_edges, ints_sides;
  std::vector<double> dinfo, dv, dp;
  for (int h = 0; h <= n; h++) {
    const bool cube = h == n;
    const int V = cube ? 8 : hull_v[h + 1] - hull_v[h], F = cube ? 6 : hull_f[h + 1] - hull_f[h];
    const int L = cube ? 24 : hull_l[h + 1] - hull_l[h];
    if (V < 4 || F < 4 || F > bm::HULL_MAX_FACES || L < 3 * F)
      return fail(ctx, BM_ERR_UNSUPPORTED, "bm_set_hulls: a hull has 4 to 64 faces of 3 or more corners");
    const double* v = cube ? CUBE_V : vertices + 3 * (size_t)hull_v[h];
    const double* p = cube ? CUBE_P : planes + 4 * (size_t)hull_f[h];
    const int32_t* lp = cube ? CUBE_L : loops + hull_l[h];
    std::vector<int32_t> st(F + 1);
    for (int f = 0; f < F; f++) st[f] = cube ? CUBE_S[f] : start[hull_f[h] + f];
    st[F] = L;
    for (int f = 0; f < F; f++) {
      if (st[f] < 0 || st[f + 1] < st[f] + 3 || st[f + 1] > L)
        return fail(ctx, BM_ERR_ARGUMENT, "bm_set_hulls: a face's loop is out of order or too short");
      if (st[f + 1] - st[f] > bm::HULL_MAX_CORNERS)
        return fail(ctx, BM_ERR_UNSUPPORTED, "bm_set_hulls: a face has more than 16 corners");
    }
    for (int i = 0; i < L; i++)
      if (lp[i] < 0 || lp[i] >= V) return fail(ctx, BM_ERR_ARGUMENT, "bm_set_hulls: a loop names no corner");
    std::vector<int32_t> edges, sides;
    double inner, reach;
    derive_hull(V, F, st.data(), lp, v, p, edges, sides, inner, reach);
    const int E = (int)edges.size() / 2;
    if (E > bm::HULL_MAX_EDGES) return fail(ctx, BM_ERR_UNSUPPORTED, "bm_set_hulls: a hull has more than 96 edges");
    info.insert(info.end(), {(int32_t)(dv.size() / 3), (int32_t)(dp.size() / 4), (int32_t)ints_start.size(),
                             (int32_t)ints_loops.size(), (int32_t)(ints_edges.size() / 2), V, F, E});
    dinfo.push_back(inner);
    dinfo.push_back(reach);
    dv.insert(dv.end(), v, v + 3 * V);
    dp.insert(dp.end(), p, p + 4 * F);
    ints_start.insert(ints_start.end(), st.begin(), st.end());
    ints_loops.insert(ints_loops.end(), lp, lp + L);
    ints_edges.insert(ints_edges.end(), edges.begin(), edges.end());
    ints_sides.insert(ints_sides.end(), sides.begin(), sides.end());
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_set_hulls: hipSetDevice");
  std::vector<int32_t> ints;
  const size_t i_start = info.size(), i_loops = i_start + ints_start.size(), i_edges = i_loops + ints_loops.size(),
               i_sides = i_edges + ints_edges.size();
  for (auto* part : {&info, &ints_start, &ints_loops, &ints_edges, &ints_sides}) ints.insert(ints.end(), part->begin(), part->end());
  std::vector<double> dbl;
  const size_t d_v = dinfo.size(), d_p = d_v + dv.size();
  for (auto* part : {&dinfo, &dv, &dp}) dbl.insert(dbl.end(), part->begin(), part->end());
  if (int rc = grow(ctx, ctx->hull_ints, ints.size() * sizeof(int32_t) + 16, "bm_set_hulls")) return rc;
  if (int rc = grow(ctx, ctx->hull_doubles, dbl.size() * sizeof(double) + 16, "bm_set_hulls")) return rc;
  HIP_OR_FAIL(ctx, hipMemcpy(ctx->hull_ints.p, ints.data(), ints.size() * sizeof(int32_t), hipMemcpyHostToDevice),
              "bm_set_hulls: copy");
  HIP_OR_FAIL(ctx, hipMemcpy(ctx->hull_doubles.p, dbl.data(), dbl.size() * sizeof(double), hipMemcpyHostToDevice),
              "bm_set_hulls: copy");
  const int32_t* I = static_cast<const int32_t*>(ctx->hull_ints.p);
  const double* D = static_cast<const double*>(ctx->hull_doubles.p);
  ctx->hulls.n = n + 1;
  ctx->hulls.cube = n;
  ctx->hulls.info = I;
  ctx->hulls.start = I + i_start;
  ctx->hulls.loops = I + i_loops;
  ctx->hulls.edges = I + i_edges;
  ctx->hulls.sides = I + i_sides;
  ctx->hulls.dinfo = D;
  ctx->hulls.vertices = D + d_v;
  ctx->hulls.planes = D + d_p;
  ctx->hull_count = n;
  return BM_OK;
}

int bm_set_launch(bm_context* ctx, int32_t shape, int32_t resident) {
  if (!ctx) return BM_ERR_ARGUMENT;
  if (shape < 0 || shape >= bm::SOFT_SHAPES) return fail(ctx, BM_ERR_ARGUMENT, "bm_set_launch: shape must be 0, 1 or 2");
  ctx->shape = shape;
  ctx->resident = resident ? 1 : 0;
  return BM_OK;
}

int bm_set_fields(bm_context* ctx, int32_t n, const int32_t* columns, const int32_t* rows, const double* spacing,
                  const int64_t* offsets, const double* heights) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || (n > 0 && (!columns || !rows || !spacing || !offsets || !heights)))
    return fail(ctx, BM_ERR_ARGUMENT, "bm_set_fields: null argument");
  std::vector<int32_t> info(2 * (size_t)n + 2);
  std::vector<long long> off((size_t)n + 1);
  std::vector<double> dinfo(3 * (size_t)n + 3);
  std::vector<bm_context::FieldShape> shapes(n);
  for (int f = 0; f < n; f++) {
    if (columns[f] < 2 || rows[f] < 2 || !(spacing[f] > 0) || offsets[f] < 0)
      return fail(ctx, BM_ERR_ARGUMENT, "bm_set_fields: a heightfield has 2 or more rows and columns, spacing above 0");
    // src/mesh.js:296-297: centred on its body, x along columns, z along rows.
    const double x0 = -((columns[f] - 1) * spacing[f]) / 2, z0 = -((rows[f] - 1) * spacing[f]) / 2;
    info[2 * f] = columns[f]; info[2 * f + 1] = rows[f];
    off[f] = offsets[f];
    dinfo[3 * f] = spacing[f]; dinfo[3 * f + 1] = x0; dinfo[3 * f + 2] = z0;
    shapes[f] = {columns[f], rows[f], spacing[f], x0, z0};
  }
  size_t total = 0;
  for (int f = 0; f < n; f++) total = std::max(total, (size_t)offsets[f] + (size_t)columns[f] * rows[f]);
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_set_fields: hipSetDevice");
  const size_t ib = info.size() * 4, ob = off.size() * 8, db = dinfo.size() * 8, hb = total * 8;
  if (int rc = grow(ctx, ctx->field_ints, ib + ob + 16, "bm_set_fields")) return rc;
  if (int rc = grow(ctx, ctx->field_doubles, db + hb + 16, "bm_set_fields")) return rc;
  char* fi = static_cast<char*>(ctx->field_ints.p);
  char* fd = static_cast<char*>(ctx->field_doubles.p);
  const size_t ob_at = (ib + 7) / 8 * 8;
  HIP_OR_FAIL(ctx, hipMemcpy(fi, info.data(), ib, hipMemcpyHostToDevice), "bm_set_fields: copy");
  HIP_OR_FAIL(ctx, hipMemcpy(fi + ob_at, off.data(), ob, hipMemcpyHostToDevice), "bm_set_fields: copy");
  HIP_OR_FAIL(ctx, hipMemcpy(fd, dinfo.data(), db, hipMemcpyHostToDevice), "bm_set_fields: copy");
  if (hb) HIP_OR_FAIL(ctx, hipMemcpy(fd + db, heights, hb, hipMemcpyHostToDevice), "bm_set_fields: copy");
  ctx->fields.n = n;
  ctx->fields.info = reinterpret_cast<const int*>(fi);
  ctx->fields.off = reinterpret_cast<const long long*>(fi + ob_at);
  ctx->fields.dinfo = reinterpret_cast<const double*>(fd);
  ctx->fields.heights = reinterpret_cast<const double*>(fd + db);
  ctx->field_shapes.swap(shapes);
  return BM_OK;
}

}  // extern "C"

namespace {

// The level pass up to its launch, for bm_level_host and bm_level_packed: the inputs checked,
// each pair's share of scratch (7 points for each triangle under its box, src/mesh.js:388-401),
// the inputs sent up through pinned staging and the kernel queued; d holds where everything is.
// The staging buffer's first in_bytes hold the inputs until the stream has sent them.
int level_launch(bm_context* ctx, const char* who, int32_t n, const int32_t* field, const double* position,
                 const double* axes, const double* shapes, const double* boxes, const double* margin,
                 const double* tolerance, int32_t cap, bm::LevelDev& d, size_t* in_bytes_out) {
  const std::string w(who);
  if (n < 0 || cap < 1) return fail(ctx, BM_ERR_ARGUMENT, w + ": n must be 0 or more, cap 1 or more");
  if (!field || !position || !axes || !shapes || !boxes || !margin || !tolerance)
    return fail(ctx, BM_ERR_ARGUMENT, w + ": null argument");
  std::vector<long long> scratch((size_t)n + 1, 0);
  for (int32_t k = 0; k < n; k++) {
    const double kind = shapes[19 * (size_t)k];
    if (!(kind == 0 || kind == 1 || kind == 2 || kind == 3))
      return fail(ctx, BM_ERR_ARGUMENT, w + ": a shape's kind must be 0, 1, 2 or 3");
    if (!hull_ok(ctx, shapes + 19 * (size_t)k))
      return fail(ctx, BM_ERR_ARGUMENT, w + ": a hull's number is not in the hull table (bm_set_hulls)");
    if (field[k] < 0 || field[k] >= (int32_t)ctx->field_shapes.size())
      return fail(ctx, BM_ERR_ARGUMENT, w + ": a field's number is not in the field table (bm_set_fields)");
    const auto& F = ctx->field_shapes[field[k]];
    const double* b = boxes + 6 * (size_t)k;
    const double i0 = host_max(0, std::floor((b[0] - F.x0) / F.spacing)), i1 = host_min(F.columns - 2, std::floor((b[3] - F.x0) / F.spacing));
    const double j0 = host_max(0, std::floor((b[2] - F.z0) / F.spacing)), j1 = host_min(F.rows - 2, std::floor((b[5] - F.z0) / F.spacing));
    long long tris = 0;
    if (i0 <= i1 && j0 <= j1) tris = 2 * (long long)(i1 - i0 + 1) * (long long)(j1 - j0 + 1);
    scratch[k + 1] = scratch[k] + 7 * tris;
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), (w + ": hipSetDevice").c_str());
  const size_t N = (size_t)n, C = (size_t)cap;
  const size_t in_d = (3 + 9 + 19 + 6 + 2) * N * 8, in_i = N * 4, in_s = (N + 1) * 8;
  const size_t in_bytes = in_d + in_s + in_i + 16;
  // Out: manifolds n, count cn, normal 3cn, points 12cn, separations 4cn, ids 4cn.
  const size_t out_d = 19 * C * N * 8, out_i = (N + 5 * C * N) * 4, out_bytes = out_d + out_i;
  if (int rc = grow(ctx, ctx->pairs_in, in_bytes, who)) return rc;
  if (int rc = grow(ctx, ctx->pairs_out, out_bytes, who)) return rc;
  const size_t points_total = (size_t)scratch[N];
  if (int rc = grow(ctx, ctx->level_scratch, points_total * (9 * 8 + 3 * 4) + 64, who)) return rc;
  char* in = static_cast<char*>(ctx->pairs_in.p);
  char* out = static_cast<char*>(ctx->pairs_out.p);
  d = bm::LevelDev{};
  d.n = n;
  d.cap = cap;
  d.position = reinterpret_cast<const double*>(in);
  d.axes = d.position + 3 * N;
  d.shapes = d.axes + 9 * N;
  d.boxes = d.shapes + 19 * N;
  d.margin = d.boxes + 6 * N;
  d.tolerance = d.margin + N;
  d.scratch = reinterpret_cast<const long long*>(in + in_d);
  d.field = reinterpret_cast<const int*>(in + in_d + in_s);
  d.normal = reinterpret_cast<double*>(out);
  d.points = d.normal + 3 * C * N;
  d.separations = d.points + 12 * C * N;
  d.manifolds = reinterpret_cast<int*>(out + out_d);
  d.count = d.manifolds + N;
  d.ids = reinterpret_cast<uint32_t*>(d.count + C * N);
  d.sd = static_cast<double*>(ctx->level_scratch.p);
  d.su = reinterpret_cast<uint32_t*>(d.sd + 9 * points_total);
  d.hulls = ctx->hulls;
  d.fields = ctx->fields;
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, in_bytes + out_bytes, who)) return rc;
  char* st = static_cast<char*>(ctx->pstage);
  double* sd = reinterpret_cast<double*>(st);
  std::memcpy(sd, position, 24 * N);
  std::memcpy(sd + 3 * N, axes, 72 * N);
  std::memcpy(sd + 12 * N, shapes, 152 * N);
  std::memcpy(sd + 31 * N, boxes, 48 * N);
  std::memcpy(sd + 37 * N, margin, 8 * N);
  std::memcpy(sd + 38 * N, tolerance, 8 * N);
  std::memcpy(st + in_d, scratch.data(), in_s);
  std::memcpy(st + in_d + in_s, field, in_i);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(in, st, in_d + in_s + in_i, hipMemcpyHostToDevice, ctx->stream), (w + ": copy").c_str());
  (void)hipGetLastError();
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), (w + ": hipEventRecord").c_str());
  bm::launch_level(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), (w + ": launch").c_str());
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), (w + ": hipEventRecord").c_str());
  *in_bytes_out = in_bytes;
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_level_host(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                  const double* shapes, const double* boxes, const double* margin, const double* tolerance, int32_t cap,
                  int32_t* manifolds, int32_t* count, double* normal, double* points, double* separations,
                  uint32_t* ids) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || cap < 1) return fail(ctx, BM_ERR_ARGUMENT, "bm_level_host: n must be 0 or more, cap 1 or more");
  if (n == 0) return BM_OK;
  if (!manifolds || !count || !normal || !points || !separations || !ids)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_host: null argument");
  bm::LevelDev d{};
  size_t in_bytes = 0;
  if (int rc = level_launch(ctx, "bm_level_host", n, field, position, axes, shapes, boxes, margin, tolerance, cap, d,
                            &in_bytes))
    return rc;
  const size_t N = (size_t)n, C = (size_t)cap;
  const size_t out_d = 19 * C * N * 8, out_i = (N + 5 * C * N) * 4, out_bytes = out_d + out_i;
  char* back = static_cast<char*>(ctx->pstage) + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, ctx->pairs_out.p, out_bytes, hipMemcpyDeviceToHost, ctx->stream),
              "bm_level_host: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_level_host: kernel");
  const double* bd = reinterpret_cast<const double*>(back);
  const int32_t* bi = reinterpret_cast<const int32_t*>(back + out_d);
  std::memcpy(manifolds, bi, 4 * N);
  // Slots a lane did not fill are zero, as the oracle and the reference leave them.
  for (size_t k = 0; k < N; k++) {
    const int got = std::min<int>(bi[k], cap);
    for (size_t j = 0; j < C; j++) {
      const size_t o = k * C + j;
      if ((int)j < got) {
        count[o] = bi[N + o];
        std::memcpy(normal + 3 * o, bd + 3 * o, 24);
        std::memcpy(points + 12 * o, bd + 3 * C * N + 12 * o, 96);
        std::memcpy(separations + 4 * o, bd + 15 * C * N + 4 * o, 32);
        std::memcpy(ids + 4 * o, bi + N + C * N + 4 * o, 16);
      } else {
        count[o] = 0;
        std::memset(normal + 3 * o, 0, 24);
        std::memset(points + 12 * o, 0, 96);
        std::memset(separations + 4 * o, 0, 32);
        std::memset(ids + 4 * o, 0, 16);
      }
    }
  }
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_level_host: hipEventElapsedTime");
  return BM_OK;
}

int bm_level_packed(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                    const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                    int32_t cap, int32_t* manifolds, int32_t* start, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || cap < 1) return fail(ctx, BM_ERR_ARGUMENT, "bm_level_packed: n must be 0 or more, cap 1 or more");
  if (!start) return fail(ctx, BM_ERR_ARGUMENT, "bm_level_packed: null argument");
  start[0] = 0;
  if (n == 0) return BM_OK;
  if (!manifolds || !count || !normal || !points || !separations || !ids)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_packed: null argument");
  bm::LevelDev d{};
  size_t in_bytes = 0;
  if (int rc = level_launch(ctx, "bm_level_packed", n, field, position, axes, shapes, boxes, margin, tolerance, cap,
                            d, &in_bytes))
    return rc;
  const size_t N = (size_t)n, R = (size_t)n * cap;
  // The packed columns, and the starts, in the scratch the narrowphase's meet pass uses for its
  // inputs (the level pass's own inputs are in pairs_in).
  const size_t pk_start = 0, pk_count = (4 * (N + 1) + 255) & ~(size_t)255, pk_normal = pk_count + ((4 * R + 255) & ~(size_t)255);
  const size_t pk_points = pk_normal + 24 * R, pk_seps = pk_points + 96 * R, pk_ids = pk_seps + 32 * R;
  const size_t pk_bytes = pk_ids + 16 * R;
  if (int rc = grow(ctx, ctx->level_packed, pk_bytes, "bm_level_packed")) return rc;
  char* pk = static_cast<char*>(ctx->level_packed.p);
  bm::LevelPacked p{};
  p.start = reinterpret_cast<int*>(pk + pk_start);
  p.count = reinterpret_cast<int*>(pk + pk_count);
  p.normal = reinterpret_cast<double*>(pk + pk_normal);
  p.points = reinterpret_cast<double*>(pk + pk_points);
  p.separations = reinterpret_cast<double*>(pk + pk_seps);
  p.ids = reinterpret_cast<uint32_t*>(pk + pk_ids);
  bm::launch_level_pack(d, p, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_level_packed: pack");
  // The counts and starts first, then only the rows found.
  char* back = static_cast<char*>(ctx->pstage) + in_bytes;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back, d.manifolds, 4 * N, hipMemcpyDeviceToHost, ctx->stream), "bm_level_packed: copy");
  HIP_OR_FAIL(ctx, hipMemcpyAsync(back + 4 * N, p.start, 4 * (N + 1), hipMemcpyDeviceToHost, ctx->stream),
              "bm_level_packed: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_level_packed: kernel");
  std::memcpy(manifolds, back, 4 * N);
  std::memcpy(start, back + 4 * N, 4 * (N + 1));
  const size_t rows = (size_t)start[N];
  if (rows > 0) {
    const struct {
      void* to;
      const void* from;
      size_t bytes;
    } cols[] = {{count, p.count, 4 * rows}, {normal, p.normal, 24 * rows}, {points, p.points, 96 * rows},
                {separations, p.separations, 32 * rows}, {ids, p.ids, 16 * rows}};
    size_t at = 0;
    for (const auto& c : cols) {
      HIP_OR_FAIL(ctx, hipMemcpyAsync(back + at, c.from, c.bytes, hipMemcpyDeviceToHost, ctx->stream),
                  "bm_level_packed: copy");
      at += (c.bytes + 255) & ~(size_t)255;
    }
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_level_packed: copy");
    std::vector<HostCopy> copies;
    at = 0;
    for (const auto& c : cols) {
      copies.push_back({c.to, back + at, c.bytes});
      at += (c.bytes + 255) & ~(size_t)255;
    }
    run_copies(copies);
  }
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_level_packed: hipEventElapsedTime");
  return BM_OK;
}

int bm_level_noalloc(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                     const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                     int32_t cap, int32_t* manifolds, int32_t* count, double* normal, double* points,
                     double* separations, uint32_t* ids) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || cap < 1) return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: n must be 0 or more, cap 1 or more");
  if (n > bm::LEVEL_NOALLOC_PAIRS)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: at most " + std::to_string(bm::LEVEL_NOALLOC_PAIRS) +
                                          " pairs a call, got " + std::to_string(n));
  if (n == 0) return BM_OK;
  if (!field || !position || !axes || !shapes || !boxes || !margin || !tolerance || !manifolds || !count || !normal ||
      !points || !separations || !ids)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: null argument");
  if ((long long)n * cap > (1LL << 31) / 16)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: n x cap manifold slots is too many");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_level_noalloc: hipSetDevice");
  bm::LevelDev d{};
  d.n = n;
  d.cap = cap;
  d.field = field;
  d.position = position;
  d.axes = axes;
  d.shapes = shapes;
  d.boxes = boxes;
  d.margin = margin;
  d.tolerance = tolerance;
  d.manifolds = manifolds;
  d.count = count;
  d.normal = normal;
  d.points = points;
  d.separations = separations;
  d.ids = ids;
  d.hulls = ctx->hulls;
  d.fields = ctx->fields;
  bm::level_on_block(d, ctx->noalloc_block);
  (void)hipGetLastError();
  // The plan checks every pair and gives each its share of the block; nothing reads a pair's
  // field, hull or points until the host has seen that every pair passed.
  bm::launch_level_plan(d, ctx->noalloc_block, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_level_noalloc: launch");
  long long status[3] = {0, 0, 0};
  HIP_OR_FAIL(ctx, hipMemcpyAsync(status, bm::level_plan_status(ctx->noalloc_block), sizeof status,
                                  hipMemcpyDeviceToHost, ctx->stream), "bm_level_noalloc: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_level_noalloc: plan");
  const int flags = (int)status[1];
  const std::string pair = "pair " + std::to_string(status[2]);
  if (flags & bm::LEVEL_PLAN_KIND) return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: " + pair + ": a shape's kind must be 0, 1, 2 or 3");
  if (flags & bm::LEVEL_PLAN_HULL)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: " + pair + ": a hull's number is not in the hull table (bm_set_hulls)");
  if (flags & bm::LEVEL_PLAN_FIELD)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: " + pair + ": a field's number is not in the field table (bm_set_fields)");
  if (flags & bm::LEVEL_PLAN_POINTS)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_level_noalloc: the boxes cover room for " + std::to_string(status[0]) +
                                          " found points (7 a triangle), more than the " +
                                          std::to_string(bm::LEVEL_NOALLOC_POINTS) + " a call holds");
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->begin, ctx->stream), "bm_level_noalloc: hipEventRecord");
  bm::launch_level(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_level_noalloc: launch");
  HIP_OR_FAIL(ctx, hipEventRecord(ctx->end, ctx->stream), "bm_level_noalloc: hipEventRecord");
  bm::launch_level_clear(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_level_noalloc: launch");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_level_noalloc: kernel");
  HIP_OR_FAIL(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->begin, ctx->end), "bm_level_noalloc: hipEventElapsedTime");
  return BM_OK;
}

int bm_set_body_version(bm_context* ctx, uint64_t version) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  ctx->body_version = version;
  return BM_OK;
}

int bm_set_device_only(bm_context* ctx, int32_t keep) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  ctx->keep_device_only = keep != 0;
  ctx->uploaded = false;
  return BM_OK;
}

int bm_set_wide(bm_context* ctx, int32_t min_manifolds) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (min_manifolds < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_set_wide: min_manifolds must be 0 or more");
  ctx->wide_min = min_manifolds;
  return BM_OK;
}

int bm_set_colour_lds(bm_context* ctx, int32_t allow) {
  if (!ctx) return BM_ERR_ARGUMENT;
  if (allow < 0 || allow > 2) return fail(ctx, BM_ERR_ARGUMENT, "bm_set_colour_lds: allow is 0, 1 or 2");
  ctx->colour_in_lds = allow;
  return BM_OK;
}

int bm_soft_step_host(bm_context* ctx, bm_columns* h) {
  if (int rc = bm_upload(ctx, h)) return rc;
  if (int rc = bm_soft_step(ctx)) return rc;
  return bm_download(ctx, h);
}

}  // extern "C"

// ---- the broadphase passes ----
namespace {

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Where each per-slot field of n bodies starts in the broadphase's body block (bytes).
struct BroadLayout {
  size_t dims, gs, reach, size, local, position, rotation, velocity, angular, type, shape, total;
  explicit BroadLayout(size_t n) {
    size_t at = 0;
    auto take = [&](size_t bytes) {
      const size_t o = at;
      at = up256(at + bytes);
      return o;
    };
    dims = take(24 * n); gs = take(8 * n); reach = take(8 * n); size = take(8 * n); local = take(48 * n);
    position = take(24 * n); rotation = take(32 * n); velocity = take(24 * n); angular = take(24 * n);
    type = take(n); shape = take(n);
    total = at;
  }
};

bm::broad::Bodies broad_bodies(bm_context* ctx) {
  const BroadLayout L(ctx->broad.n);
  char* b = static_cast<char*>(ctx->broad.bodies.p);
  bm::broad::Bodies d{};
  d.type = reinterpret_cast<const uint8_t*>(b + L.type);
  d.shape = reinterpret_cast<const uint8_t*>(b + L.shape);
  d.dims = reinterpret_cast<const double*>(b + L.dims);
  d.gravity_scale = reinterpret_cast<const double*>(b + L.gs);
  d.reach = reinterpret_cast<const double*>(b + L.reach);
  d.size = reinterpret_cast<const double*>(b + L.size);
  d.local = reinterpret_cast<const double*>(b + L.local);
  d.position = reinterpret_cast<const double*>(b + L.position);
  d.rotation = reinterpret_cast<const double*>(b + L.rotation);
  d.velocity = reinterpret_cast<const double*>(b + L.velocity);
  d.angular = reinterpret_cast<const double*>(b + L.angular);
  return d;
}

bm::broad::Tree broad_tree(bm_context* ctx, int t) {
  const bm_context::Broad& bb = ctx->broad;
  return bm::broad::Tree{bb.root[t], static_cast<const bm::broad::Node*>(bb.trees[t].p)};
}

// The leaf (n ints) and tree (n bytes) of each body, in one block.
const int* broad_leaf(bm_context* ctx) { return static_cast<const int*>(ctx->broad.leaf.p); }
const int8_t* broad_which(bm_context* ctx) {
  return reinterpret_cast<const int8_t*>(static_cast<const char*>(ctx->broad.leaf.p) + up256(4 * (size_t)ctx->broad.n));
}

// A host array going up to byte `at` of a device block.
struct Up {
  size_t at;
  const void* from;
  size_t bytes;
};

// Copies `ups` into pinned staging at their places, then to `dev` in one transfer (queued on
// the context's stream: the staging buffer is in use until the stream has run it).
int stage_up(bm_context* ctx, void* dev, const std::vector<Up>& ups, const char* who) {
  size_t total = 0;
  for (const Up& u : ups) total = std::max(total, u.at + u.bytes);
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, total, who)) return rc;
  std::vector<HostCopy> copies;
  for (const Up& u : ups) copies.push_back({static_cast<char*>(ctx->pstage) + u.at, u.from, u.bytes});
  run_copies(copies);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dev, ctx->pstage, total, hipMemcpyHostToDevice, ctx->stream), who);
  return BM_OK;
}

}  // namespace

extern "C" {

int bm_broad_set_bodies(bm_context* ctx, const bm_broad_bodies* b) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (!b || b->n < 0) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_bodies: no bodies");
  const size_t n = b->n;
  if (n > 0 && (!b->type || !b->shape || !b->dims || !b->gravity_scale || !b->reach || !b->size || !b->local ||
                !b->position || !b->rotation || !b->velocity || !b->angular))
    return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_bodies: null column");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_set_bodies: hipSetDevice");
  const BroadLayout L(n);
  if (int rc = grow(ctx, ctx->broad.bodies, L.total + 256, "bm_broad_set_bodies")) return rc;
  ctx->broad.n = b->n;
  ctx->broad.have_trees = false;  // the leaf block's size follows n
  if (n == 0) return BM_OK;
  if (int rc = stage_up(ctx, ctx->broad.bodies.p,
                        {{L.dims, b->dims, 24 * n}, {L.gs, b->gravity_scale, 8 * n},
                         {L.reach, b->reach, 8 * n}, {L.size, b->size, 8 * n},
                         {L.local, b->local, 48 * n}, {L.position, b->position, 24 * n},
                         {L.rotation, b->rotation, 32 * n}, {L.velocity, b->velocity, 24 * n},
                         {L.angular, b->angular, 24 * n}, {L.type, b->type, n},
                         {L.shape, b->shape, n}},
                        "bm_broad_set_bodies"))
    return rc;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_set_bodies: copy");
  return BM_OK;
}

int bm_broad_set_state(bm_context* ctx, int32_t n, const int32_t* slots, const double* position,
                       const double* rotation, const double* velocity, const double* angular) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || (n > 0 && (!slots || !position || !rotation || !velocity || !angular)))
    return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_state: null argument");
  if (n == 0) return BM_OK;
  for (int32_t i = 0; i < n; i++)
    if (slots[i] < 0 || slots[i] >= ctx->broad.n)
      return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_state: a slot is not a body bm_broad_set_bodies gave");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_set_state: hipSetDevice");
  const size_t N = n, rows = 13 * 8 * N, idx = up256(rows);
  if (int rc = grow(ctx, ctx->broad.rows, idx + 4 * N, "bm_broad_set_state")) return rc;
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, idx + 4 * N, "bm_broad_set_state")) return rc;
  double* r = static_cast<double*>(ctx->pstage);
  for (size_t i = 0; i < N; i++) {
    const size_t s = slots[i];
    double* o = r + 13 * i;
    for (int k = 0; k < 3; k++) o[k] = position[3 * s + k];
    for (int k = 0; k < 4; k++) o[3 + k] = rotation[4 * s + k];
    for (int k = 0; k < 3; k++) o[7 + k] = velocity[3 * s + k];
    for (int k = 0; k < 3; k++) o[10 + k] = angular[3 * s + k];
  }
  std::memcpy(static_cast<char*>(ctx->pstage) + idx, slots, 4 * N);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->broad.rows.p, ctx->pstage, idx + 4 * N, hipMemcpyHostToDevice, ctx->stream),
              "bm_broad_set_state: copy");
  const bm::broad::Bodies d = broad_bodies(ctx);
  (void)hipGetLastError();
  bm::launch_broad_scatter(n, reinterpret_cast<const int*>(static_cast<char*>(ctx->broad.rows.p) + idx),
                           static_cast<const double*>(ctx->broad.rows.p), const_cast<double*>(d.position),
                           const_cast<double*>(d.rotation), const_cast<double*>(d.velocity),
                           const_cast<double*>(d.angular), ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_broad_set_state: launch");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_set_state: scatter");
  return BM_OK;
}

int bm_broad_keep(bm_context* ctx) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (!ctx->uploaded || ctx->n_worlds != 1 || ctx->bodies != ctx->broad.n)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_keep: the last batch is not one world of the broadphase's bodies");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_keep: hipSetDevice");
  const bm::broad::Bodies d = broad_bodies(ctx);
  const size_t n = ctx->broad.n;
  const struct {
    const double* to;
    const char* column;
    size_t per;
  } moves[] = {{d.position, "position", 3}, {d.rotation, "rotation", 4}, {d.velocity, "velocity", 3},
               {d.angular, "angular", 3}};
  for (const auto& m : moves)
    HIP_OR_FAIL(ctx, hipMemcpyAsync(const_cast<double*>(m.to), ctx->buffer[column_index(m.column)], 8 * m.per * n,
                                    hipMemcpyDeviceToDevice, ctx->stream),
                "bm_broad_keep: copy");
  return BM_OK;
}

int bm_broad_set_trees(bm_context* ctx, const bm_broad_tree* trees, int32_t n, const int32_t* leaf,
                       const int8_t* which) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (!trees || !leaf || !which || n != ctx->broad.n)
    return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_trees: the trees need every body bm_broad_set_bodies gave");
  for (int t = 0; t < 2; t++) {
    const bm_broad_tree& T = trees[t];
    if (T.nodes < 0 || T.root < -1 || T.root >= T.nodes ||
        (T.nodes > 0 && (!T.box || !T.child1 || !T.child2 || !T.height || !T.item)))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_trees: a tree's nodes or root are wrong");
  }
  for (int32_t s = 0; s < n; s++)
    if (which[s] >= 0 && (which[s] > 1 || leaf[s] < 0 || leaf[s] >= trees[which[s]].nodes))
      return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_set_trees: a body's leaf is not a node of its tree");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_set_trees: hipSetDevice");
  bm_context::Broad& bb = ctx->broad;
  for (int t = 0; t < 2; t++) {
    const bm_broad_tree& T = trees[t];
    const size_t m = T.nodes;
    if (int rc = grow(ctx, bb.trees[t], sizeof(bm::broad::Node) * m + 256, "bm_broad_set_trees")) return rc;
    bb.root[t] = T.root;
    bb.nodes[t] = T.nodes;
    if (m == 0) continue;
    // Each node's box and links into one record, straight into the staging buffer.
    if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, sizeof(bm::broad::Node) * m, "bm_broad_set_trees"))
      return rc;
    bm::broad::Node* nodes = static_cast<bm::broad::Node*>(ctx->pstage);
    auto fill = [&](size_t i0, size_t i1) {
      for (size_t i = i0; i < i1; i++) {
        bm::broad::Node& o = nodes[i];
        for (int k = 0; k < 6; k++) o.box[k] = T.box[6 * i + k];
        o.child1 = T.child1[i]; o.child2 = T.child2[i]; o.height = T.height[i]; o.item = T.item[i];
      }
    };
    if (m >= 32768) {
      const int threads = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
      std::vector<std::thread> pool;
      for (int k = 1; k < threads; k++) pool.emplace_back(fill, m * k / threads, m * (k + 1) / threads);
      fill(0, m / threads);
      for (std::thread& th : pool) th.join();
    } else {
      fill(0, m);
    }
    HIP_OR_FAIL(ctx, hipMemcpyAsync(bb.trees[t].p, ctx->pstage, sizeof(bm::broad::Node) * m, hipMemcpyHostToDevice,
                                    ctx->stream),
                "bm_broad_set_trees: copy");
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_set_trees: copy");
  }
  const size_t N = n;
  if (int rc = grow(ctx, bb.leaf, up256(4 * N) + N + 256, "bm_broad_set_trees")) return rc;
  if (N > 0) {
    if (int rc = stage_up(ctx, bb.leaf.p, {{0, leaf, 4 * N}, {up256(4 * N), which, N}},
                          "bm_broad_set_trees"))
      return rc;
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_set_trees: copy");
  }
  bb.have_trees = true;
  return BM_OK;
}

int bm_broad_bounds(bm_context* ctx, int32_t n, const int32_t* list, double dt, double gravity_length, double* out,
                    uint8_t* refile) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || (n > 0 && (!list || !out || !refile))) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_bounds: null argument");
  if (!ctx->broad.have_trees) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_bounds: no trees (bm_broad_set_trees)");
  if (n == 0) return BM_OK;
  for (int32_t i = 0; i < n; i++)
    if (list[i] < 0 || list[i] >= ctx->broad.n) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_bounds: a slot is not a body");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_bounds: hipSetDevice");
  const size_t N = n, outs = 16 * 8 * N;
  bm_context::Broad& bb = ctx->broad;
  if (int rc = grow(ctx, bb.rows, 4 * N, "bm_broad_bounds")) return rc;
  if (int rc = grow(ctx, bb.out, up256(outs) + N, "bm_broad_bounds")) return rc;
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, std::max(4 * N, up256(outs) + N), "bm_broad_bounds"))
    return rc;
  std::memcpy(ctx->pstage, list, 4 * N);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(bb.rows.p, ctx->pstage, 4 * N, hipMemcpyHostToDevice, ctx->stream),
              "bm_broad_bounds: copy");
  bm::BoundsDev d{};
  d.n = n;
  d.list = static_cast<const int*>(bb.rows.p);
  d.dt = dt;
  d.gl = gravity_length;
  d.bodies = broad_bodies(ctx);
  d.moving = broad_tree(ctx, 1);
  d.leaf = broad_leaf(ctx);
  d.which = broad_which(ctx);
  d.out = static_cast<double*>(bb.out.p);
  d.refile = static_cast<uint8_t*>(bb.out.p) + up256(outs);
  (void)hipGetLastError();
  bm::launch_bounds(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_broad_bounds: launch");
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->pstage, bb.out.p, up256(outs) + N, hipMemcpyDeviceToHost, ctx->stream),
              "bm_broad_bounds: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_bounds: kernel");
  run_copies({{out, ctx->pstage, outs}, {refile, static_cast<char*>(ctx->pstage) + up256(outs), N}});
  return BM_OK;
}

int bm_broad_query(bm_context* ctx, int32_t n, const int32_t* moved, int32_t* counts, int32_t* starts, int32_t* hits,
                   int64_t capacity, int64_t* total) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || !total || capacity < 0 || (n > 0 && (!moved || !counts || !starts)) || (capacity > 0 && !hits))
    return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_query: null argument");
  if (!ctx->broad.have_trees) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_query: no trees (bm_broad_set_trees)");
  *total = 0;
  if (n == 0) return BM_OK;
  for (int32_t i = 0; i < n; i++)
    if (moved[i] < 0 || moved[i] >= ctx->broad.n) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_query: a slot is not a body");
  if (capacity > INT32_MAX) capacity = INT32_MAX;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_query: hipSetDevice");
  const size_t N = n, C = capacity;
  bm_context::Broad& bb = ctx->broad;
  if (int rc = grow(ctx, bb.rows, 4 * N, "bm_broad_query")) return rc;
  if (int rc = grow(ctx, bb.out, up256(8 * N) + 256, "bm_broad_query")) return rc;
  if (int rc = grow(ctx, bb.hits, 4 * C + 4, "bm_broad_query")) return rc;
  if (int rc = grow(ctx, bb.counters, 256, "bm_broad_query")) return rc;
  if (int rc = grow_pinned(ctx, &ctx->pstage, &ctx->pstage_cap, std::max(up256(8 * N) + 256, 4 * C + 4), "bm_broad_query"))
    return rc;
  std::memcpy(ctx->pstage, moved, 4 * N);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(bb.rows.p, ctx->pstage, 4 * N, hipMemcpyHostToDevice, ctx->stream),
              "bm_broad_query: copy");
  HIP_OR_FAIL(ctx, hipMemsetAsync(bb.counters.p, 0, 8, ctx->stream), "bm_broad_query: hipMemsetAsync");
  bm::QueryDev d{};
  d.n = n;
  d.moved = static_cast<const int*>(bb.rows.p);
  d.trees[0] = broad_tree(ctx, 0);
  d.trees[1] = broad_tree(ctx, 1);
  d.leaf = broad_leaf(ctx);
  d.which = broad_which(ctx);
  d.type = broad_bodies(ctx).type;
  d.counts = static_cast<int*>(bb.out.p);
  d.starts = static_cast<int*>(bb.out.p) + N;
  d.hits = static_cast<int*>(bb.hits.p);
  d.cap = (int)C;
  d.total = static_cast<int*>(bb.counters.p);
  d.flags = static_cast<int*>(bb.counters.p) + 1;
  (void)hipGetLastError();
  bm::launch_query(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_broad_query: launch");
  char* st = static_cast<char*>(ctx->pstage);
  HIP_OR_FAIL(ctx, hipMemcpyAsync(st, bb.out.p, 8 * N, hipMemcpyDeviceToHost, ctx->stream), "bm_broad_query: copy");
  HIP_OR_FAIL(ctx, hipMemcpyAsync(st + up256(8 * N), bb.counters.p, 8, hipMemcpyDeviceToHost, ctx->stream),
              "bm_broad_query: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_query: kernel");
  const int32_t* counters = reinterpret_cast<const int32_t*>(st + up256(8 * N));
  if (counters[1] & 1)
    return fail(ctx, BM_ERR_DEVICE, "bm_broad_query: a tree is deeper than a walk's stack holds");
  run_copies({{counts, st, 4 * N}, {starts, st + 4 * N, 4 * N}});
  *total = counters[0];
  if (*total == 0 || *total > capacity) return BM_OK;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(st, bb.hits.p, 4 * (size_t)*total, hipMemcpyDeviceToHost, ctx->stream),
              "bm_broad_query: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_query: copy");
  run_copies({{hits, st, 4 * (size_t)*total}});
  return BM_OK;
}

int bm_broad_holds(bm_context* ctx, int32_t n, const int32_t* a, const int32_t* b, uint8_t* apart) {
  if (!ctx) return BM_ERR_ARGUMENT;
  ctx->error.clear();
  if (n < 0 || (n > 0 && (!a || !b || !apart))) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_holds: null argument");
  if (!ctx->broad.have_trees) return fail(ctx, BM_ERR_ARGUMENT, "bm_broad_holds: no trees (bm_broad_set_trees)");
  if (n == 0) return BM_OK;
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device), "bm_broad_holds: hipSetDevice");
  const size_t N = n, half = up256(4 * N);
  bm_context::Broad& bb = ctx->broad;
  if (int rc = grow(ctx, bb.rows, 2 * half, "bm_broad_holds")) return rc;
  if (int rc = grow(ctx, bb.out, N, "bm_broad_holds")) return rc;
  if (int rc = stage_up(ctx, bb.rows.p, {{0, a, 4 * N}, {half, b, 4 * N}}, "bm_broad_holds"))
    return rc;
  bm::HoldsDev d{};
  d.n = n;
  d.a = static_cast<const int*>(bb.rows.p);
  d.b = reinterpret_cast<const int*>(static_cast<char*>(bb.rows.p) + half);
  d.trees[0] = broad_tree(ctx, 0);
  d.trees[1] = broad_tree(ctx, 1);
  d.leaf = broad_leaf(ctx);
  d.which = broad_which(ctx);
  d.apart = static_cast<uint8_t*>(bb.out.p);
  (void)hipGetLastError();
  bm::launch_holds(d, ctx->stream);
  HIP_OR_FAIL(ctx, hipGetLastError(), "bm_broad_holds: launch");
  // The staging buffer held the pairs going up; the stream runs in order, so the flags can
  // come back into it once the kernel has read its inputs from the device.
  HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->pstage, bb.out.p, N, hipMemcpyDeviceToHost, ctx->stream), "bm_broad_holds: copy");
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream), "bm_broad_holds: kernel");
  run_copies({{apart, ctx->pstage, N}});
  return BM_OK;
}

}  // extern "C"
