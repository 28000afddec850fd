// A whole world stepped as the reference steps it (src/world.js step, 880-1072), for
// bodies that are boxes, spheres, capsules and hulls, no joints, characters, sensors or compounds.
//
// Which parts run where. The order-defining serial bookkeeping stays on the host, in the
// reference's exact order: the AABB trees and fat boxes (src/broadphase.js, src/tree.js),
// the pair list with its swap-on-remove and awake/asleep runs and manifold cache
// (src/pairs.js), the narrow driver's visit order, reuse decisions, emit / carry /
// remember and mid-step wakes (src/contact.js:296-1036), freezing, touched and sleeping
// islands (src/world.js:1116-1343). The device does the per-pair narrowphase (collide,
// in batches between the points where a result can wake an island), the colouring and
// row layout, and the whole soft step (colour.hip, collide.hip, soft.hip). Every
// arithmetic step is the reference's, so a world's transforms are the reference's bits.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bearing_mi355x.h"
#include "tree.h"

namespace {

// An allocator whose resize() leaves new elements as it finds them, for columns every slot
// of which is written before it is read.
template <class T>
struct Unzeroed : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = Unzeroed<U>;
  };
  Unzeroed() = default;
  template <class U>
  Unzeroed(const Unzeroed<U>&) noexcept {}
  template <class U>
  void construct(U*) noexcept {}
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};

// A small pool of host threads for per-world phases and a large world's passes: run(n, f) calls f(0..n-1)
// across the threads and the caller, and returns when every call has.
class Pool {
 public:
  explicit Pool(int threads) {
    for (int t = 0; t < threads; t++) workers_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  void run(int n, const std::function<void(int)>& f) {
    if (workers_.empty() || n <= 1) {
      for (int i = 0; i < n; i++) f(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      job_ = &f;
      n_ = n;
      next_.store(0);
      busy_ = (int)workers_.size();
      ++generation_;
    }
    cv_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) f(i);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void work() {
    long seen = 0;
    for (;;) {
      const std::function<void(int)>* job;
      int n;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
        job = job_;
        n = n_;
      }
      for (int i; (i = next_.fetch_add(1)) < n;) (*job)(i);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--busy_ == 0) done_.notify_all();
      }
    }
  }
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0, busy_ = 0;
  long generation_ = 0;
  std::atomic<int> next_{0};
  bool stop_ = false;
};
constexpr int PARALLEL_VISITS = 8192;  // by default, a narrow pass this long decides its pairs on threads
constexpr int P = 4;
constexpr int DYNAMIC = 0, STATIC = 1;
constexpr int BOX = 0, SPHERE = 1, CAPSULE = 2, HULL = 3, MESH = 5, HEIGHTFIELD = 6;  // src/contact.js:114
constexpr int MOVING_TREE = 1, STATIC_TREE = 0;
constexpr double PIXEL = 1.0 / 200;      // src/contact.js:125
// Fast bodies from which a large world's catch sweeps run on its host threads.
constexpr int CATCH_PARALLEL = 256;
constexpr double FAT = 1.0 / 4;          // src/broadphase.js:32
constexpr double AHEAD = 16;             // src/broadphase.js:54
constexpr double SLEEP_TIME = 1;         // src/world.js:154
constexpr int STILL = 1, QUIET = 2, CALM = 4, HANDLED = 8, THAWING = 16;  // src/world.js:159

inline double js_max(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}
inline double js_min(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}

// src/core/math.js:223-228
void axes_of(const double* q, double* ax) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  ax[0] = 1 - 2 * (y * y + z * z); ax[1] = 2 * (x * y + z * w); ax[2] = 2 * (x * z - y * w);
  ax[3] = 2 * (x * y - z * w); ax[4] = 1 - 2 * (x * x + z * z); ax[5] = 2 * (y * z + x * w);
  ax[6] = 2 * (x * z + y * w); ax[7] = 2 * (y * z - x * w); ax[8] = 1 - 2 * (x * x + y * y);
}

// src/core/math.js:211-220 (rotate): v' = v + w t + u x t with t = 2 u x v; `inverse` turns by the conjugate.
void rotate(const double* q, double x, double y, double z, bool inverse, double* out) {
  const double sg = inverse ? -1 : 1;
  const double ux = q[0] * sg, uy = q[1] * sg, uz = q[2] * sg, w = q[3];
  const double tx = 2 * (uy * z - uz * y), ty = 2 * (uz * x - ux * z), tz = 2 * (ux * y - uy * x);
  out[0] = x + w * tx + (uy * tz - uz * ty);
  out[1] = y + w * ty + (uz * tx - ux * tz);
  out[2] = z + w * tz + (ux * ty - uy * tx);
}

// What a level pair's manifolds left behind (src/pairs.js:384-415 Memory): per manifold, its
// normal, key and what Pairs keeps for one. Only rows below count are ever read.
struct Memory {
  int count = 0;
  std::vector<double> normal, friction, slip, twist, normal_impulse;
  std::vector<uint8_t> taken, points;
  std::vector<int32_t> key;
  std::vector<uint32_t> pid;
  void ensure(int n) {
    if ((int)points.size() >= n) return;
    size_t c = points.size() < 1 ? 1 : points.size();
    while ((int)c < n) c *= 2;
    normal.resize(3 * c); friction.resize(3 * c); slip.resize(3 * c); twist.resize(3 * c);
    normal_impulse.resize(P * c); pid.resize(P * c);
    taken.resize(c); points.resize(c); key.resize(c);
  }
};

// The last step's contact events (src/events.js): which pairs began touching and which
// ended, by slot (the reference names bodies by id, generation * 2^24 + slot). A begin holds
// where the two met, the normal from a to b and the closing speed; an end holds only who.
// Cleared as a step begins (src/world.js:888); the columns keep their storage.
struct Events {
  int count = 0;
  std::vector<uint8_t> type;  // BM_EVENT_BEGIN or BM_EVENT_END
  std::vector<int32_t> a, b;
  std::vector<double> point, normal, speed;
  int add(uint8_t kind, int x, int y) {
    if (count == (int)type.size()) {
      const size_t c = std::max<size_t>(16, 2 * type.size());
      type.resize(c); a.resize(c); b.resize(c); point.resize(3 * c); normal.resize(3 * c); speed.resize(c);
    }
    const int e = count++;
    type[e] = kind; a[e] = x; b[e] = y;
    for (int k = 0; k < 3; k++) point[3 * e + k] = normal[3 * e + k] = 0;
    speed[e] = 0;
    return e;
  }
  // src/events.js:41-51
  void begin(int x, int y, double px, double py, double pz, double nx, double ny, double nz, double s) {
    const int e = add(BM_EVENT_BEGIN, x, y);
    point[3 * e] = px; point[3 * e + 1] = py; point[3 * e + 2] = pz;
    normal[3 * e] = nx; normal[3 * e + 1] = ny; normal[3 * e + 2] = nz;
    speed[e] = s;
  }
  void end(int x, int y) { add(BM_EVENT_END, x, y); }
};

// The pair list (src/pairs.js:40-378): [0, active) awake, [active, count) asleep.
struct Pairs {
  Events* events = nullptr;  // the world's, where part() writes an end
  int count = 0, active = 0;
  std::vector<int32_t> a, b, cache;
  std::vector<uint8_t> points, touching;
  std::vector<uint32_t> pid;
  std::vector<double> normal_impulse, friction, slip, twist;
  std::unordered_map<uint64_t, int> index;
  // The manifold cache (REUSE, src/contact.js:387-467), a row per cached pair.
  int cached = 0;
  std::vector<int32_t> owner;
  std::vector<uint8_t> cache_points;
  std::vector<double> cache_pose, cache_loss, cache_normal, cache_local;
  std::vector<uint32_t> cache_id;
  // Each pair's Memory (src/pairs.js:84 memory), moved and swapped with it: made when a level
  // pair first keeps something, so the other pairs (most of them) cost a null pointer.
  std::vector<std::unique_ptr<Memory>> memory;
  Memory& memory_of(int p) {
    if (!memory[p]) memory[p].reset(new Memory());
    return *memory[p];
  }
  void clear_memory(int p) {
    if (memory[p]) memory[p]->count = 0;
  }

  static uint64_t key(int x, int y) { return (uint64_t)x * 16777216ull + (uint64_t)y; }

  void grow(int n) {
    if ((int)a.size() >= n) return;
    size_t c = a.size() < 16 ? 16 : a.size();
    while ((int)c < n) c *= 2;
    a.resize(c); b.resize(c); cache.resize(c); points.resize(c); touching.resize(c); memory.resize(c);
    pid.resize(P * c); normal_impulse.resize(P * c); friction.resize(3 * c); slip.resize(3 * c); twist.resize(3 * c);
  }
  void cache_grow(int n) {
    if ((int)owner.size() >= n) return;
    size_t c = owner.size() < 16 ? 16 : owner.size();
    while ((int)c < n) c *= 2;
    owner.resize(c); cache_points.resize(c); cache_pose.resize(7 * c); cache_loss.resize(c);
    cache_normal.resize(3 * c); cache_local.resize(6 * P * c); cache_id.resize(P * c);
  }
  // src/pairs.js:243-251
  int cache_of(int p) {
    if (cache[p] != 0) return cache[p] - 1;
    const int r = cached++;
    cache_grow(cached);
    owner[r] = p;
    cache[p] = r + 1;
    return r;
  }
  // src/pairs.js:253-268
  void uncache(int p) {
    const int r = cache[p] - 1, last = --cached;
    cache[p] = 0;
    if (r == last) return;
    const int q = owner[last];
    owner[r] = q;
    cache[q] = r + 1;
    cache_points[r] = cache_points[last];
    std::memmove(&cache_pose[7 * r], &cache_pose[7 * last], 7 * sizeof(double));
    cache_loss[r] = cache_loss[last];
    std::memmove(&cache_normal[3 * r], &cache_normal[3 * last], 3 * sizeof(double));
    std::memmove(&cache_local[6 * P * r], &cache_local[6 * P * last], 6 * P * sizeof(double));
    std::memmove(&cache_id[P * r], &cache_id[P * last], P * sizeof(uint32_t));
  }
  // src/pairs.js:279-284: pair p is no longer touching; its end, if it was.
  void part(int p) {
    if (touching[p] == 0) return;
    touching[p] = 0;
    if (events) events->end(a[p], b[p]);
  }
  // src/pairs.js:234-240
  void forget(int p) {
    points[p] = 0;
    if (cache[p] != 0) uncache(p);
    clear_memory(p);
    part(p);
  }
  // src/pairs.js:303-318
  void move(int from, int to) {
    a[to] = a[from]; b[to] = b[from];
    points[to] = points[from]; touching[to] = touching[from]; cache[to] = cache[from];
    if (cache[to] != 0) owner[cache[to] - 1] = to;
    for (int k = 0; k < P; k++) { pid[P * to + k] = pid[P * from + k]; normal_impulse[P * to + k] = normal_impulse[P * from + k]; }
    for (int k = 0; k < 3; k++) {
      friction[3 * to + k] = friction[3 * from + k]; slip[3 * to + k] = slip[3 * from + k]; twist[3 * to + k] = twist[3 * from + k];
    }
    std::swap(memory[to], memory[from]);
    index[key(a[to], b[to])] = to;
  }
  // src/pairs.js:321-334
  void swap(int i, int j) {
    if (i == j) return;
    std::swap(a[i], a[j]); std::swap(b[i], b[j]);
    std::swap(points[i], points[j]); std::swap(touching[i], touching[j]); std::swap(cache[i], cache[j]);
    if (cache[i] != 0) owner[cache[i] - 1] = i;
    if (cache[j] != 0) owner[cache[j] - 1] = j;
    for (int k = 0; k < P; k++) { std::swap(pid[P * i + k], pid[P * j + k]); std::swap(normal_impulse[P * i + k], normal_impulse[P * j + k]); }
    for (int k = 0; k < 3; k++) {
      std::swap(friction[3 * i + k], friction[3 * j + k]); std::swap(slip[3 * i + k], slip[3 * j + k]);
      std::swap(twist[3 * i + k], twist[3 * j + k]);
    }
    std::swap(memory[i], memory[j]);
    index[key(a[i], b[i])] = i;
    index[key(a[j], b[j])] = j;
  }
  void wake(int p) { swap(p, active); active++; }
  void sleep(int p) { active--; swap(p, active); }
  // src/pairs.js:217-232
  void remove(int p) {
    part(p);
    if (cache[p] != 0) uncache(p);
    clear_memory(p);
    index.erase(key(a[p], b[p]));
    if (p < active) {
      const int last = --active;
      if (p != last) move(last, p);
      p = last;
    }
    const int end = --count;
    if (p != end) move(end, p);
  }
};

// A manifold as collide() or reuse leaves it (src/collide.js Manifold).
struct Manifold {
  int count = 0;
  double normal[3] = {0, 0, 0};
  double points[12] = {};
  double separations[4] = {};
  uint32_t ids[4] = {};
  double loss = 0;
};

// A row a narrow pass emitted, its columns filled when the pass stops (fill_rows): the
// manifold (a visit's rebuilt one, or a level batch's output slot), the pair's ends and
// numbers, and where its warm start comes from (the pair list, or row `kept` of a level
// pair's Memory, -1 for none).
struct RowFill {
  const Manifold* mf;  // null for a level row
  int m, p, a, b, sub, level_at, kept, shift_at;
  double tolerance, sign;
};

// A pair visited by narrow(): what find() decided for it, in the order visited.
// Most visits of a settled pile find the boxes apart, so a visit is only its decision; a kept
// manifold is rebuilt into the pass's `rebuilt` row of the same number.
struct Visit {
  int p, a, b;
  int kind;  // 0: the boxes apart (forget); 1: the kept manifold, rebuilt; 2: met on the device, batch row
             // `slot`; 3: a level pair touched on the device, level batch row `slot`
  int slot;
  bool again = false;  // prepared a second time, a body having woken since the pass began
  double tolerance;
};

// This step's manifolds in the order narrow() found them (src/contact.js Contacts).
struct Contacts {
  int count = 0;
  std::vector<int32_t> a, b, pair, sub, row, colour, order;
  std::vector<uint8_t> points, sliding;
  std::vector<double> normal, t1, t2, friction, restitution, tolerance, patch, roll_impulse1, roll_impulse2, spin_impulse,
      tangent_impulse1, tangent_impulse2, drift1, drift2, separation, point, shift, normal_impulse, max_normal_impulse,
      relative_velocity, gap, now, ra, rb, la, lb, adjusted, normal_mass, jac, block, lever, anchor_a, anchor_b, fjac,
      tangent_mass, roll_mass1, roll_mass2, spin_mass;
  std::vector<uint32_t> pid;
  void ensure(int n) {
    if ((int)a.size() >= n) return;
    size_t c = a.size() < 16 ? 16 : a.size();
    while ((int)c < n) c *= 2;
    for (auto* v : {&a, &b, &pair, &sub, &row, &colour, &order}) v->resize(c);
    points.resize(c); sliding.resize(c);
    for (auto* v : {&friction, &restitution, &tolerance, &patch, &roll_impulse1, &roll_impulse2, &spin_impulse,
                    &tangent_impulse1, &tangent_impulse2, &drift1, &drift2, &gap, &roll_mass1, &roll_mass2, &spin_mass})
      v->resize(c);
    for (auto* v : {&normal, &t1, &t2, &anchor_a, &anchor_b, &tangent_mass}) v->resize(3 * c);
    for (auto* v : {&separation, &normal_impulse, &max_normal_impulse, &relative_velocity, &now, &adjusted, &normal_mass,
                    &lever})
      v->resize(P * c);
    for (auto* v : {&point, &shift, &ra, &rb, &la, &lb}) v->resize(12 * c);
    jac.resize(48 * c); block.resize(16 * c); fjac.resize(24 * c);
    pid.resize(P * c);
  }
};

}  // namespace

// The steps a world records for the adjoint (bm_world_record): each step's solved
// manifolds held on their bodies as REUSE keeps one (src/contact.js:540-570 remember):
// each point's ends in their bodies' frames, the normal in a's frame; with the row order,
// colours, narrow order and, for each, the row whose store its pair carried.
struct Record {
  bool on = false;
  std::string broken;  // why the recorded steps cannot be held, if they cannot
  int top = 0, steps = 0;
  long catches = 0;
  double dt = 0;
  std::vector<double> initial, left, ceiling;  // the bodies at the start: p q v w; tumble state
  // rows that hold what the pairs kept before recording began: count, ids and stores
  std::vector<uint8_t> first_count;
  std::vector<uint32_t> first_pid;
  std::vector<double> first_impulse, first_friction, first_twist, first_slip;
  std::vector<int32_t> row_offset;  // steps + 1, after the first rows
  std::vector<int32_t> starts, narrow, a, b, src;
  std::vector<uint8_t> count;
  std::vector<uint32_t> pid;
  std::vector<double> local_normal, ends, shift, tolerance, patch;
  std::unordered_map<uint64_t, int32_t> last;  // a pair of bodies: the row that stored its memory last
  // this step's found manifolds, before the phase: held normal and ends, whether the pair
  // carried anything
  std::vector<double> pre_normal, pre_ends, pre_shift;
  std::vector<uint8_t> pre_had;
};

struct bm_world {
  bm_context* ctx = nullptr;
  std::string error;
  // A step that failed after it began changing the world (the bodies half moved): every
  // later step returns this code and message rather than going on from that state.
  int broken = BM_OK;
  std::string broken_msg;
  // The last step's solve: rows (points of the solved manifolds), the compulsory bytes its
  // sweeps read and write (soft_bytes), the soft-step kernel's device time.
  int64_t step_rows = 0;
  double step_bytes = 0, step_soft_ms = 0;
  double gravity[3] = {0, -9.81, 0};
  int substeps = 8;
  bool sleeping = true;
  int top = 0;
  // Body columns by slot (src/world.js:228-311), the subset this path reads and writes.
  std::vector<double> position, rotation, velocity, angular, dims, mass, inv_mass, inertia, inv_inertia, inv_i, axes,
      box, motion, start, start_q, tumbled, ceiling, gravity_scale, damping, angular_damping, size, reach, friction,
      restitution, still, rest, least;
  std::vector<uint8_t> type, shape, frozen, calm, lock, framed, awake, fast, caught;
  // A hull body's hull (its number in the table), and a framed body's offset and frame.
  std::vector<int32_t> hull;
  std::vector<double> offset, frame;
  // The hull table, as the host reads it (corners for the catch's sweeps, each hull's box
  // in its own frame and its thinnest extent); the device holds its own copy.
  struct HostHull {
    int v0, V;
    double lo[3], hi[3], size;
  };
  std::vector<HostHull> hulls;
  std::vector<double> hull_vertices;
  // A heightfield body's field (its number in the field table), and the table as the host
  // reads it (src/mesh.js:288-311: samples, the box in its own frame, its reach).
  std::vector<int32_t> field;
  struct HostField {
    int columns, rows;
    double spacing, x0, z0, lo[3], hi[3], reach;
    size_t offset;
    double at(const double* h, int i, int j) const { return h[offset + (size_t)j * columns + i]; }
  };
  std::vector<HostField> fields;
  std::vector<double> field_heights;
  std::vector<int32_t> joint_head, in_awake, awake_list, island, island_next, parent, slept;
  std::vector<float> transforms, previous;
  // moved() (src/world.js:771-803): each slot's stamp, the step it last moved in (or the
  // next step, for a body added since); the log of slots whose stamp changed other than by
  // moving awake, or that left the awake list (src/world.js:1465-1492), in stamp order;
  // entries stamped log_floor or earlier were dropped. seen marks slots already listed.
  std::vector<int64_t> moved_at, log_at;
  std::vector<int32_t> log_slots;
  std::vector<uint32_t> seen;
  int64_t log_last = 0, log_floor = 0;
  uint32_t seen_mark = 0;
  int awake_count = 0, frozen_count = 0, slept_count = 0;
  bool still_ready = false, wake_pending = false;
  double step_dt = 0;
  int step_count = 0;
  long impacts = 0, solver_steps = 0, catches = 0, catch_steps = 0;
  // Awake dynamic bodies outside their fat boxes, listed once a step when a sweep asks
  // (src/query.js:138-167); strays_at is the step they were listed in.
  std::vector<int32_t> strays;
  int strays_at = -1;
  // Broadphase (src/broadphase.js:56-209).
  bm::Tree trees[2];
  std::vector<int32_t> leaf, moved;
  std::vector<int8_t> which;
  std::vector<uint8_t> is_moved;
  Events events;
  Pairs pairs;
  Contacts c;
  // Names this world's bodies' constant columns (dims, inertias, damping, type, shape, ...): a
  // new number, unique across worlds, each time bodies are added; the device keeps those columns
  // while the number it was given stays the same (bm_set_body_version).
  uint64_t body_version = 0;
  bm_world() { pairs.events = &events; }
  // Narrowphase batches for the device: the records (filled whole, so growing them zeroes
  // nothing), then the results.
  // A meet record is its bodies' slots (batch_ia, batch_ib) and the pair's numbers; when the
  // pass goes to the device each body that meets any gives one row of 32 numbers (meet_table,
  // rows in first-met order: meet_bodies), and each pair its two rows (meet_ta, meet_tb).
  std::vector<int32_t, Unzeroed<int32_t>> batch_ia, batch_ib, meet_ta, meet_tb;
  std::vector<double, Unzeroed<double>> batch_pair, meet_table;
  std::vector<int32_t> meet_row_of, meet_bodies;  // meet_row_of: by slot, -1 when not in the table
  // The meet pass's results: count and flags by slot; the manifolds found packed, slot k's in
  // row batch_row[k] (-1 when it found none) of normal, points, seps, ids, shift and loss.
  std::vector<double> batch_normal, batch_points, batch_seps, batch_shift, batch_loss;
  std::vector<int32_t> batch_count, batch_flags, batch_row;
  std::vector<uint32_t> batch_ids;
  // Level pairs for the device touchLevel (bm_level_packed): field, the field body's place and
  // axes, the other body's shape record, the box to look in, margin and tolerance; results
  // packed by pair (pair k's manifolds from row level_start[k]), at most `level_cap` a pair.
  std::vector<int32_t, Unzeroed<int32_t>> level_field;
  std::vector<int32_t> level_manifolds, level_count, level_start;
  std::vector<double, Unzeroed<double>> level_pos, level_axes, level_shape, level_box, level_margin, level_tol;
  std::vector<double> level_normal, level_points, level_seps;
  std::vector<uint32_t> level_ids;
  int level_cap = 8;
  std::vector<Visit> visits;  // the pass's visits are the first visit_count (the vector never shrinks)
  std::vector<Manifold, Unzeroed<Manifold>> rebuilt;  // by visit: a kind-1 visit's kept manifold
  std::vector<int32_t> live;  // the visits the serial walk takes, in order (the rest done on the pool)
  std::vector<RowFill> fills;  // rows emitted since the pass last stopped, columns not yet filled
  size_t live_count = 0;
  size_t visit_count = 0;
  std::vector<uint8_t> apart;  // a large pass's pairs whose boxes no longer overlap, by slot
  // Bodies woken while a narrow pass's results are taken (their bounds changed then).
  std::vector<uint8_t> fresh;
  std::vector<int32_t> woken;
  // Where the step's narrow passes stand: the next visit to take, and the pass's first pair.
  size_t visit_at = 0;
  int narrow_from = 0;
  // A large pass decided on the pool whose first batch's results have not been looked at yet.
  bool first_results = false;
  // From how many pairs (or awake bodies, or rows) a world stepped alone takes a pass's work to
  // its host threads (bm_world_set_parallel).
  int parallel_from = PARALLEL_VISITS;
  // Within a pass, each run of the device batches appends its results: the meet results and
  // level pairs before these slots are earlier runs', still read by the visits not yet taken
  // (a visit prepared again after a wake gets a new slot past them). The records waiting for
  // the next run are those slots on.
  size_t meet_base = 0, level_base = 0, meet_found = 0;
  // For a large pass's first run: each level row's kept entry (what finish_level would match it
  // with), worked out on the pool; level_matched says these hold for visits not prepared again.
  std::vector<int32_t> level_best;
  bool level_matched = false;
  // The device context: the world's own, or its batch's (shared, and then every device call
  // made while the batch's host threads run holds device_lock).
  bool owns_ctx = true;
  std::mutex* device_lock = nullptr;
  // Host threads for a large world's narrow passes (a world stepped alone); none in a batch.
  Pool* pool = nullptr;
  std::vector<std::vector<int>> candidate_hits;  // the pooled candidate walks' overlaps
  // The broadphase's per-body and per-pair passes on the device (bm_broad_*): the way this world
  // takes them (BM_BROAD_*), and on_device for the step under way. The device holds the bodies
  // (given once: bodies_set) and their state as the last soft step left it (state), but for the
  // slots the host has changed since (dirty); the trees as of version tree_sent (each change to
  // either tree counts one in tree_version). apart_ready: w->apart holds the first pass's holds.
  struct Broad {
    int mode = BM_BROAD_AUTO;
    bool on_device = false, bodies_set = false, state = false, apart_ready = false;
    std::vector<int32_t> dirty, counts, starts, hits;
    uint64_t tree_version = 1, tree_sent = 0;
    std::vector<double> local, out;
    std::vector<uint8_t> refile;
  } broad;
  // Milliseconds spent in each part of the steps so far (bm_world_times).
  double times[BM_WORLD_TIMES] = {};
  // In a batch the device holds every world's hulls and fields in one table: this world's
  // hull h is the device's hull_base + h, its field f the device's field_base + f. The
  // world's own tables as given, for the batch to join; `joined` false until it has.
  int hull_base = 0, field_base = 0;
  bool* tables_stale = nullptr;
  std::vector<int32_t> given_hull_v, given_hull_f, given_hull_l, given_start, given_loops, given_columns, given_rows;
  std::vector<double> given_vertices, given_planes, given_spacing, given_heights;
  std::vector<int64_t> given_offsets;
  // The phase's columns as bm_upload takes them (offset tables of this world alone).
  int32_t h_body_offset[2], h_manifold_offset[2], h_pair_offset[2], h_awake_offset[2];
  int32_t h_starts[BM_STARTS], h_solved, h_substeps;
  double h_dt;
  bm_columns h;
  Record record;
  bm_columns_tap tap = nullptr;  // bm_world_tap
  void* tap_user = nullptr;
};

namespace {

int fail(bm_world* w, int code, const std::string& m) {
  w->error = m;
  return code;
}

// A step failed part-way: the world keeps the failure, and every later step returns it.
int broke(bm_world* w, int code) {
  w->broken = code;
  w->broken_msg = w->error + " (the step stopped part-way; this world takes no more steps)";
  return code;
}

// src/world.js:1917-1930 (_pose): slot s placed at its own origin and turned its own way,
// from the step's centre of mass and principal axes; the product is src/core/math.js:230-237
// (quatMultiply) of the rotation by the frame's inverse.
void pose_of(const bm_world* w, int s, double* out) {
  const double* p = &w->position[3 * s];
  const double* q = &w->rotation[4 * s];
  if (w->framed[s] == 0) {
    out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
    out[3] = q[0]; out[4] = q[1]; out[5] = q[2]; out[6] = q[3];
    return;
  }
  const double* o = &w->offset[3 * s];
  const double* f = &w->frame[4 * s];
  double v[3];
  rotate(q, o[0], o[1], o[2], false, v);
  out[0] = p[0] - v[0]; out[1] = p[1] - v[1]; out[2] = p[2] - v[2];
  const double ax = q[0], ay = q[1], az = q[2], aw = q[3];
  const double bx = -f[0], by = -f[1], bz = -f[2], bw = f[3];
  out[3] = aw * bx + ax * bw + ay * bz - az * by;
  out[4] = aw * by + ay * bw + az * bx - ax * bz;
  out[5] = aw * bz + az * bw + ax * by - ay * bx;
  out[6] = aw * bw - ax * bx - ay * by - az * bz;
}

void write_transform(bm_world* w, int s) {
  double pose[7];
  pose_of(w, s, pose);
  for (int k = 0; k < 7; k++) w->transforms[7 * s + k] = (float)pose[k];
}

// ---- bodies ----

// src/world.js:1580-1630 (_boundsOf) for a box, a sphere, a capsule or a hull.
void bounds_of(bm_world* w, int s, double dt, double gl) {
  const int s3 = 3 * s, s4 = 4 * s, a9 = 9 * s, b6 = 6 * s;
  double* ax = &w->axes[a9];
  axes_of(&w->rotation[s4], ax);
  double m = 0;
  if (w->type[s] != STATIC) {
    const double* v = &w->velocity[s3];
    const double* om = &w->angular[s3];
    const double speed = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        + (w->type[s] == DYNAMIC ? gl * std::fabs(w->gravity_scale[s]) * dt : 0)
        + std::sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]) * w->reach[s];
    m = speed * dt;
  }
  w->motion[s] = m;
  double ex, ey, ez, ox = 0, oy = 0, oz = 0;
  if (w->shape[s] == BOX || w->shape[s] == HULL || w->shape[s] >= MESH) {
    double hx = w->dims[s3] * 0.5, hy = w->dims[s3 + 1] * 0.5, hz = w->dims[s3 + 2] * 0.5;
    if (w->shape[s] != BOX) {
      // The box round its corners (or triangles) in its own frame, turned into the world's.
      const double* lo = w->shape[s] == HULL ? w->hulls[w->hull[s]].lo : w->fields[w->field[s]].lo;
      const double* hi = w->shape[s] == HULL ? w->hulls[w->hull[s]].hi : w->fields[w->field[s]].hi;
      const double cx = (lo[0] + hi[0]) * 0.5, cy = (lo[1] + hi[1]) * 0.5, cz = (lo[2] + hi[2]) * 0.5;
      hx = (hi[0] - lo[0]) * 0.5; hy = (hi[1] - lo[1]) * 0.5; hz = (hi[2] - lo[2]) * 0.5;
      ox = ax[0] * cx + ax[3] * cy + ax[6] * cz;
      oy = ax[1] * cx + ax[4] * cy + ax[7] * cz;
      oz = ax[2] * cx + ax[5] * cy + ax[8] * cz;
    }
    ex = std::fabs(ax[0]) * hx + std::fabs(ax[3]) * hy + std::fabs(ax[6]) * hz;
    ey = std::fabs(ax[1]) * hx + std::fabs(ax[4]) * hy + std::fabs(ax[7]) * hz;
    ez = std::fabs(ax[2]) * hx + std::fabs(ax[5]) * hy + std::fabs(ax[8]) * hz;
  } else if (w->shape[s] == SPHERE) {
    ex = ey = ez = w->dims[s3];
  } else {
    const double r = w->dims[s3], hl = (w->dims[s3 + 1] - 2 * r) * 0.5;
    ex = std::fabs(ax[3]) * hl + r; ey = std::fabs(ax[4]) * hl + r; ez = std::fabs(ax[5]) * hl + r;
  }
  const double grow = m + w->size[s] * PIXEL;
  const double cx = w->position[s3] + ox, cy = w->position[s3 + 1] + oy, cz = w->position[s3 + 2] + oz;
  double* box = &w->box[b6];
  box[0] = cx - ex - grow; box[1] = cy - ey - grow; box[2] = cz - ez - grow;
  box[3] = cx + ex + grow; box[4] = cy + ey + grow; box[5] = cz + ez + grow;
}

double gravity_length(const bm_world* w) {
  const double* g = w->gravity;
  return std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
}

// src/world.js:1301-1308 (_settle)
void settle(bm_world* w, int s) {
  w->still[s] = 0;
  for (int k = 0; k < 3; k++) w->rest[7 * s + k] = w->position[3 * s + k];
  for (int k = 0; k < 4; k++) w->rest[7 * s + 3 + k] = w->rotation[4 * s + k];
}

// src/world.js:1290-1296 (_awaken)
void awaken(bm_world* w, int s) {
  if (w->in_awake[s] != 0) return;
  w->awake_list[w->awake_count] = s;
  w->in_awake[s] = ++w->awake_count;
  w->awake[s] = 1;
  settle(w, s);
}

// src/world.js:1465-1492 (_log): slot s into the moved log with a stamp at least its own and
// at least the last entry's; past one entry a body, the older half goes and moved() asked
// about steps that old reads every body's stamp instead.
void log_moved(bm_world* w, int s) {
  const int64_t stamp = std::max(w->moved_at[s], w->log_last);
  w->log_slots.push_back(s);
  w->log_at.push_back(stamp);
  w->log_last = stamp;
  const int64_t n = (int64_t)w->log_slots.size();
  if (n <= w->top) return;
  const int64_t floor = w->log_at[(n - 1) >> 1];
  int64_t from = (n - 1) >> 1;
  while (from < n && w->log_at[from] <= floor) from++;
  w->log_slots.erase(w->log_slots.begin(), w->log_slots.begin() + from);
  w->log_at.erase(w->log_at.begin(), w->log_at.begin() + from);
  w->log_floor = floor;
}

// src/world.js:1311-1320 (_unawaken)
void unawaken(bm_world* w, int s) {
  const int i = w->in_awake[s] - 1;
  if (i < 0) return;
  const int last = w->awake_list[--w->awake_count];
  w->awake_list[i] = last;
  w->in_awake[last] = i + 1;
  w->in_awake[s] = 0;
  w->awake[s] = 0;
  // It may have moved in steps a later moved() asks about.
  log_moved(w, s);
}

// src/world.js:1175-1180 (_thaw)
void thaw(bm_world* w, int s, bool afresh) {
  w->frozen[s] = 0;
  w->frozen_count--;
  w->inv_mass[s] = 1 / w->mass[s];
  if (afresh) settle(w, s);
}

// src/world.js:1326-1343 (_wake)
void wake(bm_world* w, int s) {
  if (w->island[s] == 0) {
    w->calm[s] |= HANDLED;
    if (w->frozen[s] != 0) thaw(w, s, true);
    return;
  }
  const double dt = w->step_dt;
  const double gl = dt > 0 ? gravity_length(w) : 0;
  for (int b = w->island[s] - 1; b >= 0;) {
    const int next = w->island_next[b] - 1;
    w->island[b] = 0;
    w->island_next[b] = 0;
    awaken(w, b);
    if (dt > 0) bounds_of(w, b, dt, gl);
    if (!w->fresh[b]) {
      w->fresh[b] = 1;
      w->woken.push_back(b);
    }
    b = next;
  }
  w->wake_pending = true;
}

// ---- broadphase (src/broadphase.js) ----

void fat(bm_world* w, int s, int tree, double time, double* out) {
  const double* box = &w->box[6 * s];
  if (tree != MOVING_TREE) {
    for (int k = 0; k < 6; k++) out[k] = box[k];
    return;
  }
  const double m = w->size[s] * FAT;
  for (int k = 0; k < 3; k++) {
    const double d = w->velocity[3 * s + k] * time;
    out[k] = box[k] - m + (d < 0 ? d : 0);
    out[k + 3] = box[k + 3] + m + (d > 0 ? d : 0);
  }
}

void mark(bm_world* w, int s) {
  if (w->is_moved[s]) return;
  w->is_moved[s] = 1;
  w->moved.push_back(s);
}

void broadphase_add(bm_world* w, int s) {
  const int t = w->type[s] == STATIC ? STATIC_TREE : MOVING_TREE;
  double out[6];
  fat(w, s, t, 0, out);
  w->leaf[s] = w->trees[t].insert(s, out);
  w->which[s] = (int8_t)t;
  w->broad.tree_version++;
  mark(w, s);
}

// src/broadphase.js:119-124 (settle: a body fallen asleep keeps a fat box with no lead)
void broadphase_settle(bm_world* w, int s) {
  if (w->which[s] != MOVING_TREE) return;
  double out[6];
  fat(w, s, MOVING_TREE, 0, out);
  w->trees[MOVING_TREE].move(w->leaf[s], out);
  w->broad.tree_version++;
}

// src/broadphase.js:140-159
void broadphase_update(bm_world* w, double dt) {
  bm::Tree& tree = w->trees[MOVING_TREE];
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (w->which[s] != MOVING_TREE) continue;
    const double* fb = &tree.box[6 * w->leaf[s]];
    const double* b = &w->box[6 * s];
    if (b[0] >= fb[0] && b[1] >= fb[1] && b[2] >= fb[2] && b[3] <= fb[3] && b[4] <= fb[4] && b[5] <= fb[5]) continue;
    double out[6];
    fat(w, s, MOVING_TREE, AHEAD * dt, out);
    tree.move(w->leaf[s], out);
    w->broad.tree_version++;
    mark(w, s);
  }
}

bool overlap(bm_world* w, int a, int b) {
  const double* fa = &w->trees[w->which[a]].box[6 * w->leaf[a]];
  const double* fb = &w->trees[w->which[b]].box[6 * w->leaf[b]];
  return fa[0] <= fb[3] && fa[3] >= fb[0] && fa[1] <= fb[4] && fa[4] >= fb[1] && fa[2] <= fb[5] && fa[5] >= fb[2];
}

// src/pairs.js:180-214 (found), for bodies with no sensors, layers or joints.
void found(bm_world* w, int x, int y) {
  if (w->type[x] != DYNAMIC && w->type[y] != DYNAMIC) return;
  if (w->type[y] == STATIC) wake(w, x);
  else if (w->type[x] == STATIC) wake(w, y);
  const int a = x < y ? x : y, b = x < y ? y : x;
  Pairs& pr = w->pairs;
  if (pr.index.count(Pairs::key(a, b))) return;
  const int p = pr.count;
  pr.grow(p + 1);
  pr.a[p] = a; pr.b[p] = b; pr.points[p] = 0; pr.touching[p] = 0; pr.cache[p] = 0;
  pr.clear_memory(p);
  pr.index[Pairs::key(a, b)] = p;
  pr.count++;
  pr.wake(p);
}

// A large world's candidate walks run on its pool: each walk only reads the trees (found()
// changes bodies' boxes and the pairs, never a tree's nodes), so every moved body's overlaps
// are gathered at once, then handed to found() in the order the serial loop takes them.
constexpr int CANDIDATES_PARALLEL = 512;

bool candidates_pooled(bm_world* w) {
  const int n = (int)w->moved.size();
  if (!w->pool || n < CANDIDATES_PARALLEL) return false;
  const int chunk = 128, chunks = (n + chunk - 1) / chunk;
  auto& hits = w->candidate_hits;  // per chunk: each body's overlaps, then -1
  if ((int)hits.size() < chunks) hits.resize(chunks);
  w->pool->run(chunks, [&](int c) {
    thread_local std::vector<int> stack;
    std::vector<int>& out = hits[c];
    out.clear();
    for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++) {
      const int a = w->moved[i];
      const int t = w->which[a];
      if (t >= 0) {
        double q[6];
        std::memcpy(q, &w->trees[t].box[6 * w->leaf[a]], sizeof q);
        auto visit = [&](int b) {
          if (b != a) out.push_back(b);
        };
        if (w->type[a] != STATIC) w->trees[STATIC_TREE].query(q, visit, stack);
        w->trees[MOVING_TREE].query(q, visit, stack);
      }
      out.push_back(-1);
    }
  });
  for (int c = 0; c < chunks; c++) {
    size_t k = 0;
    const std::vector<int>& out = hits[c];
    for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++) {
      const int a = w->moved[i];
      w->is_moved[a] = 0;
      for (; out[k] >= 0; k++) found(w, a, out[k]);
      k++;
    }
  }
  w->moved.clear();
  return true;
}

// src/broadphase.js:161-179
void candidates(bm_world* w) {
  if (candidates_pooled(w)) return;
  for (size_t i = 0; i < w->moved.size(); i++) {
    const int a = w->moved[i];
    w->is_moved[a] = 0;
    const int t = w->which[a];
    if (t < 0) continue;
    double q[6];
    std::memcpy(q, &w->trees[t].box[6 * w->leaf[a]], sizeof q);
    auto visit = [&](int b) {
      if (b != a) found(w, a, b);
    };
    if (w->type[a] != STATIC) w->trees[STATIC_TREE].query(q, visit);
    w->trees[MOVING_TREE].query(q, visit);
  }
  w->moved.clear();
}

// ---- the broadphase's passes on the device (bm_broad_*) ----

// A world takes them on the device when it has its own context and, in BM_BROAD_AUTO, at
// least BROAD_AWAKE awake bodies; then the walks when at least BROAD_WALKS bodies were marked
// and the holds when a pass starts over at least BROAD_HOLDS pairs (in BM_BROAD_DEVICE, any).
constexpr int BROAD_AWAKE = 4096, BROAD_WALKS = 256, BROAD_HOLDS = 16384;

bool broad_on(const bm_world* w) {
  if (!w->owns_ctx || w->broad.mode == BM_BROAD_HOST) return false;
  return w->broad.mode == BM_BROAD_DEVICE || w->awake_count >= BROAD_AWAKE;
}

int broad_fail(bm_world* w, int rc) { return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx)); }

// Every body's constants and state, once (and again after bodies, hulls or fields change).
int broad_give_bodies(bm_world* w) {
  const int n = w->top;
  std::vector<double>& local = w->broad.local;
  local.assign(6 * (size_t)n, 0);
  for (int s = 0; s < n; s++) {
    if (w->shape[s] != HULL && w->shape[s] < MESH) continue;
    const double* lo = w->shape[s] == HULL ? w->hulls[w->hull[s]].lo : w->fields[w->field[s]].lo;
    const double* hi = w->shape[s] == HULL ? w->hulls[w->hull[s]].hi : w->fields[w->field[s]].hi;
    double* l = &local[6 * (size_t)s];
    for (int k = 0; k < 3; k++) {
      l[k] = (lo[k] + hi[k]) * 0.5;
      l[3 + k] = (hi[k] - lo[k]) * 0.5;
    }
  }
  bm_broad_bodies b{};
  b.n = n;
  b.type = w->type.data(); b.shape = w->shape.data(); b.dims = w->dims.data();
  b.gravity_scale = w->gravity_scale.data(); b.reach = w->reach.data(); b.size = w->size.data();
  b.local = local.data(); b.position = w->position.data(); b.rotation = w->rotation.data();
  b.velocity = w->velocity.data(); b.angular = w->angular.data();
  if (int rc = bm_broad_set_bodies(w->ctx, &b)) return broad_fail(w, rc);
  w->broad.bodies_set = w->broad.state = true;
  w->broad.dirty.clear();
  w->broad.tree_sent = 0;  // the leaf table is sized by the bodies
  return BM_OK;
}

// The trees, when either changed since they were last sent.
int broad_give_trees(bm_world* w) {
  if (w->broad.tree_sent == w->broad.tree_version) return BM_OK;
  bm_broad_tree t[2];
  for (int k = 0; k < 2; k++) {
    const bm::Tree& T = w->trees[k];
    t[k].nodes = (int32_t)T.parent.size();
    t[k].root = T.root;
    t[k].box = T.box.data(); t[k].child1 = T.child1.data(); t[k].child2 = T.child2.data();
    t[k].height = T.height.data(); t[k].item = T.item.data();
  }
  if (int rc = bm_broad_set_trees(w->ctx, t, w->top, w->leaf.data(), w->which.data())) return broad_fail(w, rc);
  w->broad.tree_sent = w->broad.tree_version;
  return BM_OK;
}

// src/world.js:1574-1578 (_bounds) and src/broadphase.js:140-159 (update): every awake body's
// bounds and fat-box test on the device, then the re-filing in the awake list's order (a
// re-filed leaf moves no other leaf, so each body's test holds however many went before it).
int broad_update(bm_world* w, double dt, double gl) {
  if (!w->broad.bodies_set || !w->broad.state) {
    if (int rc = broad_give_bodies(w)) return rc;
  } else if (!w->broad.dirty.empty()) {
    std::vector<int32_t>& d = w->broad.dirty;
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    if (int rc = bm_broad_set_state(w->ctx, (int32_t)d.size(), d.data(), w->position.data(), w->rotation.data(),
                                    w->velocity.data(), w->angular.data()))
      return broad_fail(w, rc);
    d.clear();
  }
  if (int rc = broad_give_trees(w)) return rc;
  const int n = w->awake_count;
  std::vector<double>& out = w->broad.out;
  std::vector<uint8_t>& refile = w->broad.refile;
  if (out.size() < 16 * (size_t)n) out.resize(16 * (size_t)n);
  if (refile.size() < (size_t)n) refile.resize(n);
  if (int rc = bm_broad_bounds(w->ctx, n, w->awake_list.data(), dt, gl, out.data(), refile.data()))
    return broad_fail(w, rc);
  // The results are field-major: field k of row i at out[k * n + i].
  const int32_t* list = w->awake_list.data();
  auto take = [&](int i0, int i1) {
    for (int i = i0; i < i1; i++) {
      const size_t s = list[i];
      for (int k = 0; k < 9; k++) w->axes[9 * s + k] = out[(size_t)k * n + i];
      w->motion[s] = out[(size_t)9 * n + i];
      for (int k = 0; k < 6; k++) w->box[6 * s + k] = out[(size_t)(10 + k) * n + i];
    }
  };
  if (w->pool && n >= 8192) {
    const int chunk = 2048, chunks = (n + chunk - 1) / chunk;
    w->pool->run(chunks, [&](int c) { take(c * chunk, std::min(n, (c + 1) * chunk)); });
  } else {
    take(0, n);
  }
  bm::Tree& tree = w->trees[MOVING_TREE];
  for (int i = 0; i < n; i++) {
    if (!refile[i]) continue;
    const int s = list[i];
    double f[6];
    fat(w, s, MOVING_TREE, AHEAD * dt, f);
    tree.move(w->leaf[s], f);
    w->broad.tree_version++;
    mark(w, s);
  }
  return BM_OK;
}

// src/broadphase.js:161-179 (candidates) with the walks on the device, their bodies handed to
// found() in the walks' order; done false when too few bodies were marked to send.
int broad_candidates(bm_world* w, bool& done) {
  done = false;
  const int n = (int)w->moved.size();
  if (n == 0 || (w->broad.mode != BM_BROAD_DEVICE && n < BROAD_WALKS)) return BM_OK;
  if (int rc = broad_give_trees(w)) return rc;
  auto& b = w->broad;
  if (b.counts.size() < (size_t)n) {
    b.counts.resize(n);
    b.starts.resize(n);
  }
  if (b.hits.size() < 8 * (size_t)n) b.hits.resize(8 * (size_t)n);
  int64_t total = 0;
  for (int round = 0; round < 2; round++) {
    if (int rc = bm_broad_query(w->ctx, n, w->moved.data(), b.counts.data(), b.starts.data(), b.hits.data(),
                                (int64_t)b.hits.size(), &total))
      return broad_fail(w, rc);
    if (total <= (int64_t)b.hits.size()) break;
    b.hits.resize((size_t)total);  // the walks again, with room for every body they meet
  }
  for (int i = 0; i < n; i++) {
    const int a = w->moved[i];
    w->is_moved[a] = 0;
    for (int k = b.starts[i], e = b.starts[i] + b.counts[i]; k < e; k++) found(w, a, b.hits[k]);
  }
  w->moved.clear();
  done = true;
  return BM_OK;
}

// src/pairs.js:340-356 (holds) for a pass over pairs [from, active) on the device, into
// w->apart, for narrow_begin to take.
int broad_holds(bm_world* w, int from) {
  Pairs& pr = w->pairs;
  const int m = pr.active - from;
  if (m == 0 || (w->broad.mode != BM_BROAD_DEVICE && m < BROAD_HOLDS)) return BM_OK;
  if (int rc = broad_give_trees(w)) return rc;
  if ((int)w->apart.size() < pr.active) w->apart.resize(pr.active);
  if (int rc = bm_broad_holds(w->ctx, m, pr.a.data() + from, pr.b.data() + from, w->apart.data() + from))
    return broad_fail(w, rc);
  w->broad.apart_ready = true;
  return BM_OK;
}

// ---- narrow driver (src/contact.js:296-1036) ----

// src/contact.js:686-693 (moving)
double moving(const bm_world* w, int a, int b) {
  const double* v = w->velocity.data();
  const double* om = w->angular.data();
  const double* g = w->gravity;
  const int a3 = 3 * a, b3 = 3 * b;
  const double ga = w->type[a] == DYNAMIC ? w->gravity_scale[a] : 0, gb = w->type[b] == DYNAMIC ? w->gravity_scale[b] : 0;
  const double dx = v[b3] - v[a3], dy = v[b3 + 1] - v[a3 + 1], dz = v[b3 + 2] - v[a3 + 2];
  const double gs = gb - ga, gl = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) * std::fabs(gs);
  const double turning = std::sqrt(om[a3] * om[a3] + om[a3 + 1] * om[a3 + 1] + om[a3 + 2] * om[a3 + 2]) * w->reach[a]
      + std::sqrt(om[b3] * om[b3] + om[b3 + 1] * om[b3 + 1] + om[b3 + 2] * om[b3 + 2]) * w->reach[b];
  return (std::sqrt(dx * dx + dy * dy + dz * dz) + 0.5 * gl * w->step_dt + turning) * w->step_dt;
}

// src/contact.js:609-622 (relative): the smaller body's pose in the larger's frame, into pose; returns the smaller.
int relative(const bm_world* w, int a, int b, double* pose) {
  const int h = w->reach[a] >= w->reach[b] ? a : b, g = h == a ? b : a;
  const double* pos = w->position.data();
  const double* A = w->axes.data();
  const double* q = w->rotation.data();
  const int h3 = 3 * h, g3 = 3 * g, h9 = 9 * h, h4 = 4 * h, g4 = 4 * g;
  const double dx = pos[g3] - pos[h3], dy = pos[g3 + 1] - pos[h3 + 1], dz = pos[g3 + 2] - pos[h3 + 2];
  pose[0] = A[h9] * dx + A[h9 + 1] * dy + A[h9 + 2] * dz;
  pose[1] = A[h9 + 3] * dx + A[h9 + 4] * dy + A[h9 + 5] * dz;
  pose[2] = A[h9 + 6] * dx + A[h9 + 7] * dy + A[h9 + 8] * dz;
  const double ax = -q[h4], ay = -q[h4 + 1], az = -q[h4 + 2], aw = q[h4 + 3];
  const double bx = q[g4], by = q[g4 + 1], bz = q[g4 + 2], bw = q[g4 + 3];
  pose[3] = aw * bx + ax * bw + ay * bz - az * by;
  pose[4] = aw * by + ay * bw + az * bx - ax * bz;
  pose[5] = aw * bz + az * bw + ax * by - ay * bx;
  pose[6] = aw * bw - ax * bx - ay * by - az * bz;
  return g;
}

// src/contact.js:518-537 (keeps)
bool keeps(const bm_world* w, int p, int a, int b, double tolerance) {
  const Pairs& pr = w->pairs;
  const int r = pr.cache[p] - 1;
  const double* pose = &pr.cache_pose[7 * r];
  double now[7];
  const int g = relative(w, a, b, now);
  const double ex = now[0] - pose[0], ey = now[1] - pose[1], ez = now[2] - pose[2];
  const double cx = -pose[3], cy = -pose[4], cz = -pose[5], cw = pose[6];
  const double rx = now[3], ry = now[4], rz = now[5], rw = now[6];
  const double ux = cw * rx + cx * rw + cy * rz - cz * ry;
  const double uy = cw * ry + cy * rw + cz * rx - cx * rz;
  const double uz = cw * rz + cz * rw + cx * ry - cy * rx;
  const double drift = std::sqrt(ex * ex + ey * ey + ez * ez) + 2 * std::sqrt(ux * ux + uy * uy + uz * uz) * w->reach[g];
  if (pr.cache_loss[r] + 2 * drift > tolerance) return false;
  if (w->joint_head[a] != 0 || w->joint_head[b] != 0) return false;
  return pr.touching[p] == 1 || drift + moving(w, a, b) <= tolerance;
}

// src/contact.js:573-576 (rounded): a ball's or a capsule's radius, which its kept core leaves out.
double rounded(const bm_world* w, int s) { return w->shape[s] == SPHERE || w->shape[s] == CAPSULE ? w->dims[3 * s] : 0; }

double least(const Manifold& m) {
  double s = INFINITY;
  for (int k = 0; k < m.count; k++)
    if (m.separations[k] < s) s = m.separations[k];
  return s;
}

// src/contact.js:468-510 (reused)
bool reused(const bm_world* w, int p, int a, int b, double tolerance, double margin, Manifold& m) {
  const Pairs& pr = w->pairs;
  if (pr.cache[p] == 0 || !keeps(w, p, a, b, tolerance)) return false;
  const int r = pr.cache[p] - 1, count = pr.cache_points[r];
  const double* pos = w->position.data();
  const double* A = w->axes.data();
  const int a3 = 3 * a, b3 = 3 * b, a9 = 9 * a, b9 = 9 * b;
  const double* ln = &pr.cache_normal[3 * r];
  const double lx = ln[0], ly = ln[1], lz = ln[2];
  const double nx = A[a9] * lx + A[a9 + 3] * ly + A[a9 + 6] * lz;
  const double ny = A[a9 + 1] * lx + A[a9 + 4] * ly + A[a9 + 7] * lz;
  const double nz = A[a9 + 2] * lx + A[a9 + 5] * ly + A[a9 + 8] * lz;
  m.normal[0] = nx; m.normal[1] = ny; m.normal[2] = nz;
  const double ra = rounded(w, a), rb = -rounded(w, b);
  m.count = count;
  for (int k = 0; k < count; k++) {
    const double* L = &pr.cache_local[6 * (P * r + k)];
    const double ax = L[0], ay = L[1], az = L[2], bx = L[3], by = L[4], bz = L[5];
    const double pax = pos[a3] + A[a9] * ax + A[a9 + 3] * ay + A[a9 + 6] * az + nx * ra;
    const double pay = pos[a3 + 1] + A[a9 + 1] * ax + A[a9 + 4] * ay + A[a9 + 7] * az + ny * ra;
    const double paz = pos[a3 + 2] + A[a9 + 2] * ax + A[a9 + 5] * ay + A[a9 + 8] * az + nz * ra;
    const double pbx = pos[b3] + A[b9] * bx + A[b9 + 3] * by + A[b9 + 6] * bz + nx * rb;
    const double pby = pos[b3 + 1] + A[b9 + 1] * bx + A[b9 + 4] * by + A[b9 + 7] * bz + ny * rb;
    const double pbz = pos[b3 + 2] + A[b9 + 2] * bx + A[b9 + 5] * by + A[b9 + 8] * bz + nz * rb;
    const double s = (pbx - pax) * nx + (pby - pay) * ny + (pbz - paz) * nz, half = 0.5 * s;
    m.separations[k] = s;
    if (rb != 0) {
      m.points[3 * k] = pbx - nx * half; m.points[3 * k + 1] = pby - ny * half; m.points[3 * k + 2] = pbz - nz * half;
    } else if (ra != 0) {
      m.points[3 * k] = pax + nx * half; m.points[3 * k + 1] = pay + ny * half; m.points[3 * k + 2] = paz + nz * half;
    } else {
      m.points[3 * k] = (pax + pbx) * 0.5; m.points[3 * k + 1] = (pay + pby) * 0.5; m.points[3 * k + 2] = (paz + pbz) * 0.5;
    }
    m.ids[k] = pr.cache_id[P * r + k];
  }
  if (least(m) > margin) return false;
  return true;
}

// src/contact.js:578-586 (cored): a ball's kept end is its centre, a capsule's the nearest
// point of its axis.
void cored(const bm_world* w, int s, double* local) {
  if (w->shape[s] == SPHERE) { local[0] = 0; local[1] = 0; local[2] = 0; return; }
  if (w->shape[s] != CAPSULE) return;
  const double half = (w->dims[3 * s + 1] - 2 * w->dims[3 * s]) * 0.5, y = local[1];
  local[0] = 0; local[1] = y < -half ? -half : y > half ? half : y; local[2] = 0;
}

// src/contact.js:540-570 (remember), for a manifold not found at the moment of impact.
void remember(bm_world* w, int p, int a, int b, const Manifold& m) {
  Pairs& pr = w->pairs;
  const int count = m.count, r = pr.cache_of(p);
  pr.cache_points[r] = (uint8_t)count;
  double pose[7];
  relative(w, a, b, pose);
  for (int k = 0; k < 7; k++) pr.cache_pose[7 * r + k] = pose[k];
  pr.cache_loss[r] = m.loss;
  const double* pos = w->position.data();
  const double* A = w->axes.data();
  const int a3 = 3 * a, b3 = 3 * b, a9 = 9 * a, b9 = 9 * b;
  const double nx = m.normal[0], ny = m.normal[1], nz = m.normal[2];
  double* ln = &pr.cache_normal[3 * r];
  ln[0] = A[a9] * nx + A[a9 + 1] * ny + A[a9 + 2] * nz;
  ln[1] = A[a9 + 3] * nx + A[a9 + 4] * ny + A[a9 + 5] * nz;
  ln[2] = A[a9 + 6] * nx + A[a9 + 7] * ny + A[a9 + 8] * nz;
  for (int k = 0; k < count; k++) {
    double* L = &pr.cache_local[6 * (P * r + k)];
    const double half = 0.5 * m.separations[k];
    const double px = m.points[3 * k], py = m.points[3 * k + 1], pz = m.points[3 * k + 2];
    const double ax = px - nx * half - pos[a3], ay = py - ny * half - pos[a3 + 1], az = pz - nz * half - pos[a3 + 2];
    const double bx = px + nx * half - pos[b3], by = py + ny * half - pos[b3 + 1], bz = pz + nz * half - pos[b3 + 2];
    L[0] = A[a9] * ax + A[a9 + 1] * ay + A[a9 + 2] * az;
    L[1] = A[a9 + 3] * ax + A[a9 + 4] * ay + A[a9 + 5] * az;
    L[2] = A[a9 + 6] * ax + A[a9 + 7] * ay + A[a9 + 8] * az;
    L[3] = A[b9] * bx + A[b9 + 1] * by + A[b9 + 2] * bz;
    L[4] = A[b9 + 3] * bx + A[b9 + 4] * by + A[b9 + 5] * bz;
    L[5] = A[b9 + 6] * bx + A[b9 + 7] * by + A[b9 + 8] * bz;
    cored(w, a, L);
    cored(w, b, L + 3);
    pr.cache_id[P * r + k] = m.ids[k];
  }
}

// src/contact.js:1283-1292 (basis)
void basis(Contacts& c, int m) {
  const double nx = c.normal[3 * m], ny = c.normal[3 * m + 1], nz = c.normal[3 * m + 2];
  const double ax = std::fabs(nx), ay = std::fabs(ny), az = std::fabs(nz);
  double tx, ty, tz;
  if (ax <= ay && ax <= az) { tx = 0; ty = nz; tz = -ny; }
  else if (ay <= az) { tx = -nz; ty = 0; tz = nx; }
  else { tx = ny; ty = -nx; tz = 0; }
  const double l = 1 / std::sqrt(tx * tx + ty * ty + tz * tz);
  tx *= l; ty *= l; tz *= l;
  c.t1[3 * m] = tx; c.t1[3 * m + 1] = ty; c.t1[3 * m + 2] = tz;
  c.t2[3 * m] = ny * tz - nz * ty; c.t2[3 * m + 1] = nz * tx - nx * tz; c.t2[3 * m + 2] = nx * ty - ny * tx;
}

// src/contact.js:889-921 (emit), sign 1, no sub-manifold; `shift` (3 a point) for a
// manifold met at the moment of impact, else null.
// In two halves: the row's number, taken in the pass's order (emit_slot), and its columns
// (emit_fill), which read only the manifold and the bodies' materials and write only row m.
int emit_slot(bm_world* w) {
  Contacts& c = w->c;
  const int m = c.count++;
  c.ensure(c.count);
  return m;
}
void emit_fill(bm_world* w, int m, int p, int a, int b, const Manifold& mf, double tolerance, const double* shift,
               int sub, double sign) {
  Contacts& c = w->c;
  c.a[m] = a; c.b[m] = b; c.pair[m] = p; c.sub[m] = sub;
  c.points[m] = (uint8_t)mf.count;
  for (int k = 0; k < 3; k++) c.normal[3 * m + k] = sign * mf.normal[k];
  c.friction[m] = std::sqrt(w->friction[a] * w->friction[b]);
  c.restitution[m] = js_max(w->restitution[a], w->restitution[b]);
  c.tolerance[m] = tolerance;
  c.patch[m] = PIXEL * (w->type[a] != DYNAMIC ? w->size[b] : w->type[b] != DYNAMIC ? w->size[a]
                                                                                   : js_min(w->size[a], w->size[b]));
  c.roll_impulse1[m] = 0; c.roll_impulse2[m] = 0; c.spin_impulse[m] = 0;
  c.tangent_impulse1[m] = 0; c.tangent_impulse2[m] = 0; c.drift1[m] = 0; c.drift2[m] = 0;
  for (int k = 0; k < mf.count; k++) {
    const int i = m * P + k;
    c.pid[i] = mf.ids[k];
    c.separation[i] = mf.separations[k];
    for (int j = 0; j < 3; j++) c.point[3 * i + j] = mf.points[3 * k + j];
    for (int j = 0; j < 3; j++) c.shift[3 * i + j] = shift ? shift[3 * k + j] : 0;
    c.normal_impulse[i] = 0;
  }
  basis(c, m);
}

// src/contact.js:999-1036 (carry) from the pair list, or from a row of a level pair's Memory.
template <class Kept>
void carry_from(bm_world* w, const Kept& pr, int p, int m) {
  Contacts& c = w->c;
  const int remembered = pr.points[p];
  if (remembered == 0) return;
  const int m3 = 3 * m;
  const double t1x = c.t1[m3], t1y = c.t1[m3 + 1], t1z = c.t1[m3 + 2];
  const double t2x = c.t2[m3], t2y = c.t2[m3 + 1], t2z = c.t2[m3 + 2];
  const int count = c.points[m], p3 = 3 * p;
  const double lx = pr.twist[p3], ly = pr.twist[p3 + 1], lz = pr.twist[p3 + 2];
  if (count <= 2) {
    c.roll_impulse1[m] = lx * t1x + ly * t1y + lz * t1z;
    c.roll_impulse2[m] = lx * t2x + ly * t2y + lz * t2z;
  }
  c.spin_impulse[m] = lx * c.normal[m3] + ly * c.normal[m3 + 1] + lz * c.normal[m3 + 2];
  const double fx = pr.friction[p3], fy = pr.friction[p3 + 1], fz = pr.friction[p3 + 2];
  c.tangent_impulse1[m] = fx * t1x + fy * t1y + fz * t1z;
  c.tangent_impulse2[m] = fx * t2x + fy * t2y + fz * t2z;
  bool same = false;
  for (int k = 0; k < count; k++) {
    const int i = m * P + k;
    const uint32_t id = c.pid[i];
    for (int j = 0; j < remembered; j++) {
      const int o = p * P + j;
      if (pr.pid[o] != id) continue;
      c.normal_impulse[i] = pr.normal_impulse[o];
      same = true;
      break;
    }
  }
  if (same) {
    const double sx = pr.slip[p3], sy = pr.slip[p3 + 1], sz = pr.slip[p3 + 2];
    c.drift1[m] = sx * t1x + sy * t1y + sz * t1z;
    c.drift2[m] = sx * t2x + sy * t2y + sz * t2z;
  }
}

void carry(bm_world* w, int p, int m) { carry_from(w, w->pairs, p, m); }

// A level batch's packed output row o as a manifold.
void level_manifold(const bm_world* w, size_t o, Manifold& mf) {
  mf.count = w->level_count[o];
  for (int i = 0; i < 3; i++) mf.normal[i] = w->level_normal[3 * o + i];
  for (int i = 0; i < 12; i++) mf.points[i] = w->level_points[12 * o + i];
  for (int i = 0; i < 4; i++) mf.separations[i] = w->level_seps[4 * o + i];
  for (int i = 0; i < 4; i++) mf.ids[i] = w->level_ids[4 * o + i];
}

// One emitted row's columns and warm start (src/contact.js:889-921 emit, 999-1036 carry).
void fill_row(bm_world* w, const RowFill& f) {
  if (f.mf) {
    const double* shift = f.shift_at >= 0 ? &w->batch_shift[12 * (size_t)w->batch_row[f.shift_at]] : nullptr;
    emit_fill(w, f.m, f.p, f.a, f.b, *f.mf, f.tolerance, shift, f.sub, f.sign);
    carry(w, f.p, f.m);
    return;
  }
  Manifold mf;
  level_manifold(w, (size_t)f.level_at, mf);
  emit_fill(w, f.m, f.p, f.a, f.b, mf, f.tolerance, nullptr, f.sub, f.sign);
  if (f.kept >= 0) carry_from(w, w->pairs.memory_of(f.p), f.kept, f.m);
}

// The rows emitted since the pass last stopped: each writes only its own row and reads what
// no visit changes while the pass is taken (the batches' outputs, the bodies' materials, the
// pair's own memory), so a large set is filled on the pool in any order.
constexpr int PARALLEL_FILLS = 4096;
void fill_rows(bm_world* w) {
  std::vector<RowFill>& fills = w->fills;
  const int n = (int)fills.size();
  if (w->pool && n >= PARALLEL_FILLS) {
    const int chunk = 1024, chunks = (n + chunk - 1) / chunk;
    w->pool->run(chunks, [&](int c) {
      for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++) fill_row(w, fills[i]);
    });
  } else {
    for (const RowFill& f : fills) fill_row(w, f);
  }
  fills.clear();
}

// A body record for the device meet (src/contact.js:1244-1276 load, and what moving()
// and src/impact.js start read): the shape (kind, centre 3, axes 9, half 3, weight 3),
// rotation 4, velocity 3, spin 3, gravity scale (0 unless dynamic), reach, flags
// (1 dynamic, 2 spin capped at pi/dt: dynamic and not a plain sphere).
void load(const bm_world* w, int s, double* r) {
  r[0] = w->shape[s] == BOX ? 2.0 : w->shape[s] == CAPSULE ? 1.0 : w->shape[s] == HULL ? 3.0 : 0.0;
  for (int k = 0; k < 3; k++) r[1 + k] = w->position[3 * s + k];
  for (int k = 0; k < 9; k++) r[4 + k] = w->axes[9 * s + k];
  if (w->shape[s] == HULL) {
    r[13] = w->hull_base + w->hull[s]; r[14] = 0; r[15] = 0;
  } else if (w->shape[s] == BOX) {
    for (int k = 0; k < 3; k++) r[13 + k] = w->dims[3 * s + k] * 0.5;
  } else if (w->shape[s] == CAPSULE) {
    r[13] = w->dims[3 * s]; r[14] = (w->dims[3 * s + 1] - 2 * w->dims[3 * s]) * 0.5; r[15] = 0;
  } else {
    r[13] = w->dims[3 * s]; r[14] = 0; r[15] = 0;
  }
  const double scale = w->type[s] == DYNAMIC ? w->gravity_scale[s] : 0;
  for (int k = 0; k < 3; k++) r[16 + k] = w->gravity[k] * scale;
  for (int k = 0; k < 4; k++) r[19 + k] = w->rotation[4 * s + k];
  for (int k = 0; k < 3; k++) r[23 + k] = w->velocity[3 * s + k];
  for (int k = 0; k < 3; k++) r[26 + k] = w->angular[3 * s + k];
  r[29] = scale;
  r[30] = w->reach[s];
  const bool dynamic = w->type[s] == DYNAMIC;
  r[31] = (dynamic ? 1 : 0) | (dynamic && (w->shape[s] != SPHERE || w->framed[s] == 1) ? 2 : 0);
}

// Runs the level batch gathered so far on the device (bm_level_host) and clears it; a pair
// touched along more directions than the slots run again with room for all of them.
int run_level_batch(bm_world* w) {
  const int n = (int)w->level_field.size();
  if (n == 0) return BM_OK;
  // This run's pairs and rows go after the pass's earlier runs'.
  const size_t k0 = w->level_base, r0 = k0 > 0 ? (size_t)w->level_start[k0] : 0;
  for (;;) {
    const size_t slots = r0 + (size_t)n * w->level_cap;
    w->level_manifolds.resize(k0 + n);
    w->level_start.resize(k0 + n + 1);
    w->level_count.resize(slots);
    w->level_normal.resize(3 * slots);
    w->level_points.resize(12 * slots);
    w->level_seps.resize(4 * slots);
    w->level_ids.resize(4 * slots);
    const int rc = bm_level_packed(w->ctx, n, w->level_field.data(), w->level_pos.data(), w->level_axes.data(),
                                   w->level_shape.data(), w->level_box.data(), w->level_margin.data(),
                                   w->level_tol.data(), w->level_cap, w->level_manifolds.data() + k0,
                                   w->level_start.data() + k0, w->level_count.data() + r0,
                                   w->level_normal.data() + 3 * r0, w->level_points.data() + 12 * r0,
                                   w->level_seps.data() + 4 * r0, w->level_ids.data() + 4 * r0);
    if (rc != BM_OK) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
    int most = 0;
    for (int k = 0; k < n; k++) most = std::max(most, (int)w->level_manifolds[k0 + k]);
    if (most <= w->level_cap) break;
    w->level_cap = most;
  }
  for (int k = 0; k <= n; k++) w->level_start[k0 + k] += (int32_t)r0;
  w->level_base = k0 + n;
  for (auto* v : {&w->level_pos, &w->level_axes, &w->level_shape, &w->level_box, &w->level_margin, &w->level_tol})
    v->clear();
  w->level_field.clear();
  return BM_OK;
}

// The rows of the pass's n meet records: each body once, in the order the records first name
// it (meet_bodies), and each record's two rows (meet_ta, meet_tb). Returns the bodies.
int meet_rows(bm_world* w, int n) {
  std::vector<int32_t>& of = w->meet_row_of;
  if ((int)of.size() < w->top) of.resize(w->top, -1);
  std::vector<int32_t>& list = w->meet_bodies;
  list.clear();
  w->meet_ta.resize(n);
  w->meet_tb.resize(n);
  auto row = [&](int s) {
    if (of[s] < 0) {
      of[s] = (int32_t)list.size();
      list.push_back(s);
    }
    return of[s];
  };
  for (int k = 0; k < n; k++) {
    w->meet_ta[k] = row(w->batch_ia[k]);
    w->meet_tb[k] = row(w->batch_ib[k]);
  }
  for (int s : list) of[s] = -1;
  return (int)list.size();
}

// The bodies' rows of the meet table into out (32 numbers a row), on the pool when many.
void meet_table_fill(bm_world* w, double* out) {
  const std::vector<int32_t>& list = w->meet_bodies;
  const int nb = (int)list.size();
  auto fill = [&](int i0, int i1) {
    for (int i = i0; i < i1; i++) load(w, list[i], out + 32 * (size_t)i);
  };
  if (w->pool && nb >= 8192) {
    const int chunk = 2048, chunks = (nb + chunk - 1) / chunk;
    w->pool->run(chunks, [&](int c) { fill(c * chunk, std::min(nb, (c + 1) * chunk)); });
  } else {
    fill(0, nb);
  }
}

// Runs the narrowphase batch gathered so far on the device (bm_meet_indexed) and clears it.
int run_batch(bm_world* w) {
  if (int rc = run_level_batch(w)) return rc;
  const int n = (int)w->batch_pair.size() / 6;
  if (n == 0) return BM_OK;
  const int nb = meet_rows(w, n);
  w->meet_table.resize(32 * (size_t)nb);
  meet_table_fill(w, w->meet_table.data());
  // This run's results go after the pass's earlier runs', its rows after theirs.
  const size_t k = w->meet_base, e = k + n, r = w->meet_found, re = r + n;
  w->batch_count.resize(e);
  w->batch_flags.resize(e);
  w->batch_row.resize(e);
  w->batch_normal.resize(3 * re);
  w->batch_points.resize(12 * re);
  w->batch_seps.resize(4 * re);
  w->batch_ids.resize(4 * re);
  w->batch_shift.resize(12 * re);
  w->batch_loss.resize(re);
  const int rc = bm_meet_packed(w->ctx, nb, w->meet_table.data(), n, w->meet_ta.data(), w->meet_tb.data(),
                                w->batch_pair.data(), w->batch_count.data() + k, w->batch_flags.data() + k,
                                w->batch_row.data() + k, w->batch_normal.data() + 3 * r,
                                w->batch_points.data() + 12 * r, w->batch_seps.data() + 4 * r,
                                w->batch_ids.data() + 4 * r, w->batch_shift.data() + 12 * r, w->batch_loss.data() + r);
  w->batch_ia.clear(); w->batch_ib.clear(); w->batch_pair.clear();
  if (rc != BM_OK) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
  size_t found = 0;
  for (size_t j = k; j < e; j++)
    if (w->batch_row[j] >= 0) {
      w->batch_row[j] += (int32_t)r;
      found++;
    }
  w->meet_base = e;
  w->meet_found = r + found;
  return BM_OK;
}

// src/contact.js:832-856 (level) up to touchLevel: the other body's box, grown by its motion,
// in the field's frame and grown by the field's own motion, written into the level batch at
// the visit's slot with the other body's shape.
void fill_level(bm_world* w, const Visit& v) {
  const int a = v.a, b = v.b;
  const bool turned = w->shape[a] >= MESH;
  const int g = turned ? a : b, o = turned ? b : a;
  const double tolerance = v.tolerance;
  const double* reach = w->motion.data();
  const double margin = reach[a] + reach[b] + tolerance;
  const double* box = w->box.data();
  const double* A = w->axes.data();
  const double* pos = w->position.data();
  const int o6 = 6 * o, g3 = 3 * g, g9 = 9 * g;
  const double cx = (box[o6] + box[o6 + 3]) * 0.5 - pos[g3], cy = (box[o6 + 1] + box[o6 + 4]) * 0.5 - pos[g3 + 1];
  const double cz = (box[o6 + 2] + box[o6 + 5]) * 0.5 - pos[g3 + 2];
  const double hx = (box[o6 + 3] - box[o6]) * 0.5, hy = (box[o6 + 4] - box[o6 + 1]) * 0.5, hz = (box[o6 + 5] - box[o6 + 2]) * 0.5;
  const double lx = A[g9] * cx + A[g9 + 1] * cy + A[g9 + 2] * cz;
  const double ly = A[g9 + 3] * cx + A[g9 + 4] * cy + A[g9 + 5] * cz;
  const double lz = A[g9 + 6] * cx + A[g9 + 7] * cy + A[g9 + 8] * cz;
  const double ex = std::fabs(A[g9]) * hx + std::fabs(A[g9 + 1]) * hy + std::fabs(A[g9 + 2]) * hz + reach[g];
  const double ey = std::fabs(A[g9 + 3]) * hx + std::fabs(A[g9 + 4]) * hy + std::fabs(A[g9 + 5]) * hz + reach[g];
  const double ez = std::fabs(A[g9 + 6]) * hx + std::fabs(A[g9 + 7]) * hy + std::fabs(A[g9 + 8]) * hz + reach[g];
  const size_t k = v.slot - w->level_base;
  w->level_field[k] = w->field_base + w->field[g];
  for (int i = 0; i < 3; i++) w->level_pos[3 * k + i] = pos[g3 + i];
  for (int i = 0; i < 9; i++) w->level_axes[9 * k + i] = A[g9 + i];
  double r[32];
  load(w, o, r);
  for (int i = 0; i < 19; i++) w->level_shape[19 * k + i] = r[i];
  const double lb[6] = {lx - ex, ly - ey, lz - ez, lx + ex, ly + ey, lz + ez};
  for (int i = 0; i < 6; i++) w->level_box[6 * k + i] = lb[i];
  w->level_margin[k] = margin;
  w->level_tol[k] = tolerance;
}

// The rest of src/contact.js:857-881 (level): wake, then each direction the device found as a
// manifold of its own (sub 0, 1, ...; normals from a to b), each started from the remembered
// manifold of the pair nearest its direction; a pair that touched nothing forgets.
// A level pair's rows matched to what the pair kept, in order (src/contact.js:857-880): each
// row takes the untaken entry of no key whose normal is nearest its own (the largest positive
// dot product), or none. Reads and writes only the pair's own Memory.
void match_level(bm_world* w, const Visit& v, int32_t* best_of) {
  const int a = v.a, p = v.p, k = v.slot;
  const double sign = w->shape[a] >= MESH ? -1 : 1;
  const int n = w->level_manifolds[k];
  Memory& kept = w->pairs.memory_of(p);
  for (int j = 0; j < kept.count; j++) kept.taken[j] = 0;
  for (int q = 0; q < n; q++) {
    const size_t o = (size_t)w->level_start[k] + q;
    // The row's normal as emit_fill writes it, for the match against what the pair kept.
    const double nx = sign * w->level_normal[3 * o], ny = sign * w->level_normal[3 * o + 1];
    const double nz = sign * w->level_normal[3 * o + 2];
    int best = -1;
    double most = 0;
    for (int j = 0; j < kept.count; j++) {
      if (kept.taken[j] || kept.key[j] != 0) continue;
      const double d = kept.normal[3 * j] * nx + kept.normal[3 * j + 1] * ny + kept.normal[3 * j + 2] * nz;
      if (d > most) { most = d; best = j; }
    }
    if (best >= 0) kept.taken[best] = 1;
    best_of[q] = best;
  }
}

int finish_level(bm_world* w, Visit& v) {
  const int a = v.a, b = v.b, p = v.p, k = v.slot;
  const bool turned = w->shape[a] >= MESH;
  const int n = w->level_manifolds[k];
  if (n > 0 && w->awake[a] != w->awake[b]) wake(w, w->awake[a] ? b : a);
  const double sign = turned ? -1 : 1;
  // Matched on the pool when the pass's first run came back, unless prepared again since.
  const size_t o0 = (size_t)w->level_start[k];
  int32_t own[64];
  const int32_t* best = nullptr;
  if (w->level_matched && !v.again) {
    best = &w->level_best[o0];
  } else if (n <= 64) {
    match_level(w, v, own);
    best = own;
  } else {
    w->level_best.resize(std::max(w->level_best.size(), o0 + n));
    match_level(w, v, &w->level_best[o0]);
    best = &w->level_best[o0];
  }
  int sub = 0;
  for (int q = 0; q < n; q++) {
    const int m = emit_slot(w);
    w->fills.push_back(RowFill{nullptr, m, p, a, b, sub++, (int)(o0 + q), best[q], -1, v.tolerance, sign});
  }
  if (sub == 0) w->pairs.forget(p);
  return BM_OK;
}

// find()'s decision for one visit (src/contact.js:334-354), reading only: the boxes apart
// (0), the kept manifold rebuilt (1, REUSE, into m), a meet on the device (2), a level pair (3).
void decide_visit(const bm_world* w, Visit& v, Manifold& m) {
  const int a = v.a, b = v.b;
  v.kind = 0;
  v.slot = -1;
  const double* ba = &w->box[6 * a];
  const double* bb = &w->box[6 * b];
  if (ba[0] > bb[3] || bb[0] > ba[3] || ba[1] > bb[4] || bb[1] > ba[4] || ba[2] > bb[5] || bb[2] > ba[5]) return;
  if (w->shape[a] >= MESH || w->shape[b] >= MESH) {
    // A field has no thickness: the tolerance is the other body's (src/contact.js:836).
    v.kind = 3;
    v.tolerance = w->size[w->shape[a] >= MESH ? b : a] * PIXEL;
    return;
  }
  const double margin = w->motion[a] + w->motion[b] + v.tolerance;
  v.kind = reused(w, v.p, a, b, v.tolerance, margin, m) ? 1 : 2;
}

// The meet record of a visit of kind 2 at its slot: both bodies and the pair's numbers. The
// bodies' own records are taken when the pass goes to the device: nothing they hold (pose,
// axes, velocities) changes while a pass is prepared, a wake changing only a body's motion,
// which is the pair's.
void fill_meet(bm_world* w, const Visit& v) {
  const size_t k = v.slot - w->meet_base;
  w->batch_ia[k] = v.a;
  w->batch_ib[k] = v.b;
  const double margin = w->motion[v.a] + w->motion[v.b] + v.tolerance;
  const double pair[6] = {margin, v.tolerance, w->step_dt, w->gravity[0], w->gravity[1], w->gravity[2]};
  for (int i = 0; i < 6; i++) w->batch_pair[6 * k + i] = pair[i];
}

// Room in the device batches for `meets` more meet records and `levels` more level records.
void grow_batches(bm_world* w, size_t meets, size_t levels) {
  const size_t m = w->batch_pair.size() / 6 + meets;
  w->batch_ia.resize(m); w->batch_ib.resize(m); w->batch_pair.resize(6 * m);
  const size_t l = w->level_field.size() + levels;
  w->level_field.resize(l); w->level_pos.resize(3 * l); w->level_axes.resize(9 * l); w->level_shape.resize(19 * l);
  w->level_box.resize(6 * l); w->level_margin.resize(l); w->level_tol.resize(l);
}

// find() up to the narrowphase (src/contact.js:334-354) for one visit: the box test, the
// kept manifold (REUSE), a meet record appended to the device batch, or a level pair's.
void prepare_visit(bm_world* w, Visit& v, Manifold& m) {
  decide_visit(w, v, m);
  if (v.kind == 2) {
    v.slot = (int)(w->meet_base + w->batch_pair.size() / 6);
    grow_batches(w, 1, 0);
    fill_meet(w, v);
  } else if (v.kind == 3) {
    v.slot = (int)(w->level_base + w->level_field.size());
    grow_batches(w, 0, 1);
    fill_level(w, v);
  }
}

// The rest of find() and narrow()'s own steps for one visit, in the reference's order:
// src/contact.js:339-363 (forget, or remember what meet() found) and 318-322 (wake,
// emit, carry). Reads the device batch's results for a met pair.
int finish_visit(bm_world* w, Visit& v, Manifold& m) {
  if (v.kind == 0) {
    w->pairs.forget(v.p);
    return BM_OK;
  }
  if (v.kind == 3) return finish_level(w, v);
  int shift_at = -1;
  if (v.kind == 2) {
    const int k = v.slot;
    m.count = w->batch_count[k];
    if (m.count < 0) return fail(w, BM_ERR_UNSUPPORTED, "step: a pair of shapes the device narrowphase does not handle");
    if (w->batch_flags[k] & 2) w->impacts++;
    if (m.count == 0) {
      w->pairs.forget(v.p);
      return BM_OK;
    }
    const size_t r = w->batch_row[k];
    m.loss = w->batch_loss[r];
    for (int i = 0; i < 3; i++) m.normal[i] = w->batch_normal[3 * r + i];
    for (int i = 0; i < 12; i++) m.points[i] = w->batch_points[12 * r + i];
    for (int i = 0; i < 4; i++) m.separations[i] = w->batch_seps[4 * r + i];
    for (int i = 0; i < 4; i++) m.ids[i] = w->batch_ids[4 * r + i];
    if (w->batch_flags[k] & 1) {
      // A manifold met at the moment of impact is not kept (src/contact.js:541).
      if (w->pairs.cache[v.p] != 0) w->pairs.uncache(v.p);
      shift_at = k;
    } else {
      remember(w, v.p, v.a, v.b, m);
    }
  }
  if (w->awake[v.a] != w->awake[v.b]) wake(w, w->awake[v.a] ? v.b : v.a);
  w->fills.push_back(RowFill{&m, emit_slot(w), v.p, v.a, v.b, -1, -1, -1, shift_at, v.tolerance, 1});
  return BM_OK;
}

// src/contact.js:296-327 (narrow), from pair `from` on. Pairs that no longer overlap go as
// the reference drops them; every other pair's narrowphase runs in one device batch; then
// the results are taken in the pairs' order. A pair whose manifold wakes a sleeping island
// changes the bounds of the bodies it wakes, so a later pair with a body woken that way
// is looked at again from where things stand then, as the reference, which visits pairs
// one by one, would see it.
//
// The pass is split where it needs the device: narrow_begin() prepares every pair (records for
// the device batch), narrow_resume() takes the results in order and stops, leaving a record
// behind, where a pair must be looked at again; the caller runs the batch between the two.
void narrow_begin(bm_world* w, int from) {
  Pairs& pr = w->pairs;
  // The last pass's runs are all taken: this one's results start afresh.
  w->meet_base = 0;
  w->level_base = 0;
  w->meet_found = 0;
  w->level_matched = false;
  std::vector<Visit>& visits = w->visits;
  auto visit_of = [&](int p) {
    const int a = pr.a[p], b = pr.b[p];
    return Visit{p, a, b, 0, -1, false, js_min(w->size[a], w->size[b]) * PIXEL};
  };
  const bool ready = w->broad.apart_ready;
  w->broad.apart_ready = false;
  if (ready || (w->pool && pr.active - from >= w->parallel_from)) {
    // Each overlap reads only the trees, so a large pass tests them on the device or the pool,
    // then removes in the serial loop's order (a removal moves the last active pair, and its
    // flag, into the slot); the pairs kept are then slots from..active in order, one visit each.
    const int m = pr.active - from, chunk = 4096, chunks = (m + chunk - 1) / chunk;
    std::vector<uint8_t>& apart = w->apart;
    if ((int)apart.size() < pr.active) apart.resize(pr.active);
    if (!ready)
      w->pool->run(chunks, [&](int c) {
        for (int p = from + c * chunk, e = std::min(pr.active, from + (c + 1) * chunk); p < e; p++)
          apart[p] = !overlap(w, pr.a[p], pr.b[p]);
      });
    // Only the slots whose pair is apart, found on the pool, in rising order: a slot that is not
    // apart stays so (a removal moves the last active pair, and its flag, into the removed
    // slot only), so the serial loop acts at these slots alone.
    std::vector<std::vector<int32_t>> found_per((size_t)chunks);
    auto find = [&](int c) {
      std::vector<int32_t>& f = found_per[c];
      f.clear();
      for (int p = from + c * chunk, e = std::min(pr.active, from + (c + 1) * chunk); p < e; p++)
        if (apart[p]) f.push_back(p);
    };
    if (w->pool) w->pool->run(chunks, find);
    else
      for (int c = 0; c < chunks; c++) find(c);
    for (const std::vector<int32_t>& f : found_per)
      for (const int p : f)
        while (p < pr.active && apart[p]) {
          apart[p] = apart[pr.active - 1];
          pr.remove(p);
        }
    const int kept = pr.active - from, kept_chunks = (kept + chunk - 1) / chunk;
    if ((int)visits.size() < kept) visits.resize(kept);  // every visit is written below
    auto visits_of = [&](int c) {
      for (int i = c * chunk, e = std::min(kept, (c + 1) * chunk); i < e; i++) visits[i] = visit_of(from + i);
    };
    if (w->pool) w->pool->run(kept_chunks, visits_of);
    else
      for (int c = 0; c < kept_chunks; c++) visits_of(c);
    w->visit_count = kept;
  } else {
    size_t k = 0;
    for (int p = from; p < pr.active;) {
      if (!overlap(w, pr.a[p], pr.b[p])) {
        pr.remove(p);
        continue;
      }
      if (k < visits.size()) visits[k] = visit_of(p);
      else visits.push_back(visit_of(p));
      k++;
      p++;
    }
    w->visit_count = k;
  }
  // The pairs are where they will stay for the pass: each one's decision reads only, so a
  // large pass decides on the pool's threads, then takes its slots in order and fills them.
  const int n = (int)w->visit_count;
  if (w->rebuilt.size() < w->visits.size()) w->rebuilt.resize(w->visits.size());
  Manifold* rebuilt = w->rebuilt.data();
  if (!w->pool && w->owns_ctx && n >= w->parallel_from) {
    // A world stepped alone that grew large: its own threads, up to 16.
    const int threads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    w->pool = new Pool(threads - 1);
  }
  if (w->pool && n >= w->parallel_from) {
    const int chunk = 1024, chunks = (n + chunk - 1) / chunk;
    // Each chunk's meets and level pairs counted as it decides, so their slots (in visit
    // order) are known per chunk and taken on the pool below.
    std::vector<size_t> meets_at((size_t)chunks + 1, 0), levels_at((size_t)chunks + 1, 0);
    w->pool->run(chunks, [&](int c) {
      size_t nm = 0, nl = 0;
      for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++) {
        decide_visit(w, visits[i], rebuilt[i]);
        nm += visits[i].kind == 2;
        nl += visits[i].kind == 3;
      }
      meets_at[c + 1] = nm;
      levels_at[c + 1] = nl;
    });
    meets_at[0] = w->meet_base + w->batch_pair.size() / 6;
    levels_at[0] = w->level_base + w->level_field.size();
    const size_t m0 = meets_at[0], l0 = levels_at[0];
    for (int c = 0; c < chunks; c++) {
      meets_at[c + 1] += meets_at[c];
      levels_at[c + 1] += levels_at[c];
    }
    grow_batches(w, meets_at[chunks] - m0, levels_at[chunks] - l0);
    // A pair whose boxes are apart forgets what it kept; when it keeps no cache row, was not
    // touching (so it writes no end event) and neither body can wake while the pass is taken
    // (so the pair is never looked at again), that touches only the pair's own entries, in any
    // order: those are forgotten here, and the serial walk takes the other visits, in order.
    std::vector<int> kept_per((size_t)chunks + 1, 0);
    const uint8_t* type = w->type.data();
    const uint8_t* awake = w->awake.data();
    auto can_wake = [&](int b) { return type[b] == DYNAMIC && awake[b] == 0; };
    w->pool->run(chunks, [&](int c) {
      int kept_here = 0;
      size_t meet = meets_at[c], level = levels_at[c];
      for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++) {
        Visit& v = visits[i];
        if (v.kind == 2) {
          v.slot = (int)meet++;
          fill_meet(w, v);
        } else if (v.kind == 3) {
          v.slot = (int)level++;
          fill_level(w, v);
        } else if (v.kind == 0 && pr.cache[v.p] == 0 && pr.touching[v.p] == 0 && !can_wake(v.a) && !can_wake(v.b)) {
          pr.points[v.p] = 0;
          pr.clear_memory(v.p);
          pr.part(v.p);
          v.kind = 4;
          continue;
        }
        kept_here++;
      }
      kept_per[c + 1] = kept_here;
    });
    for (int c = 0; c < chunks; c++) kept_per[c + 1] += kept_per[c];
    if (w->live.size() < (size_t)kept_per[chunks]) w->live.resize(kept_per[chunks]);
    int32_t* live = w->live.data();
    w->pool->run(chunks, [&](int c) {
      int at = kept_per[c];
      for (int i = c * chunk, e = std::min(n, (c + 1) * chunk); i < e; i++)
        if (visits[i].kind != 4) live[at++] = i;
    });
    w->live_count = kept_per[chunks];
    w->first_results = true;
  } else {
    for (int i = 0; i < n; i++) prepare_visit(w, visits[i], rebuilt[i]);
    if (w->live.size() < (size_t)n) w->live.resize(n);
    for (int i = 0; i < n; i++) w->live[i] = i;
    w->live_count = n;
  }
  // Prepared after every wake so far: nothing is fresh for this pass yet.
  for (int s : w->woken) w->fresh[s] = 0;
  w->woken.clear();
  w->visit_at = 0;
}

// Takes the pass's results in order; true when it stopped for the device (a pair prepared
// again left a record in the batch), false when the pass is done.
// A large pass's first batch back: the meets that found nothing for a pair that keeps no
// cache row, was not touching (so it writes no end event) and neither of whose bodies can wake
// while the pass is taken (so it is never prepared again) forget their pair, as finish_visit
// would, on the pool and in any order: each touches only its own pair's entries, and the impact
// count is a sum. The serial walk then takes the other visits, in order. On a pile at rest these
// are most of the walk (a pair whose boxes touch but whose shapes do not).
void drop_empty_meets(bm_world* w) {
  const int n = (int)w->live_count, chunk = 2048, chunks = (n + chunk - 1) / chunk;
  Pairs& pr = w->pairs;
  const uint8_t* type = w->type.data();
  const uint8_t* awake = w->awake.data();
  auto can_wake = [&](int b) { return type[b] == DYNAMIC && awake[b] == 0; };
  std::vector<int> kept_per((size_t)chunks + 1, 0), impacts_per((size_t)chunks, 0);
  int32_t* live = w->live.data();
  Visit* visits = w->visits.data();
  w->pool->run(chunks, [&](int c) {
    int kept = 0, impacts = 0;
    for (int j = c * chunk, e = std::min(n, (c + 1) * chunk); j < e; j++) {
      Visit& v = visits[live[j]];
      if (v.kind == 2 && w->batch_count[v.slot] == 0 && pr.cache[v.p] == 0 && pr.touching[v.p] == 0 &&
          !can_wake(v.a) && !can_wake(v.b)) {
        if (w->batch_flags[v.slot] & 2) impacts++;
        pr.points[v.p] = 0;
        pr.clear_memory(v.p);
        v.kind = 4;
        continue;
      }
      kept++;
    }
    kept_per[c + 1] = kept;
    impacts_per[c] = impacts;
  });
  for (int c = 0; c < chunks; c++) {
    kept_per[c + 1] += kept_per[c];
    w->impacts += impacts_per[c];
  }
  // The rest keep their order: each chunk's survivors move down to where the chunks before end.
  std::vector<int32_t> rest((size_t)kept_per[chunks]);
  w->pool->run(chunks, [&](int c) {
    int at = kept_per[c];
    for (int j = c * chunk, e = std::min(n, (c + 1) * chunk); j < e; j++)
      if (visits[live[j]].kind != 4) rest[at++] = live[j];
  });
  std::copy(rest.begin(), rest.end(), live);
  w->live_count = kept_per[chunks];
}

// A large pass's first run back: every level pair's rows matched to what it kept, on the pool
// (each pair reads and writes only its own Memory, and no visit before it in the walk changes
// that), so the serial walk only takes their rows' places.
void match_levels(bm_world* w) {
  const size_t rows = w->level_base > 0 ? (size_t)w->level_start[w->level_base] : 0;
  if (w->level_best.size() < rows) w->level_best.resize(rows);
  const int n = (int)w->live_count, chunk = 2048, chunks = (n + chunk - 1) / chunk;
  const int32_t* live = w->live.data();
  Visit* visits = w->visits.data();
  w->pool->run(chunks, [&](int c) {
    for (int j = c * chunk, e = std::min(n, (c + 1) * chunk); j < e; j++) {
      const Visit& v = visits[live[j]];
      if (v.kind == 3) match_level(w, v, &w->level_best[(size_t)w->level_start[v.slot]]);
    }
  });
  w->level_matched = true;
}

bool narrow_resume(bm_world* w, int& rc) {
  if (w->first_results) {
    w->first_results = false;
    if (w->pool && w->visit_at == 0) {
      match_levels(w);
      drop_empty_meets(w);
    }
  }
  std::vector<Visit>& visits = w->visits;
  constexpr size_t AHEAD_VISITS = 16;  // visits asked for before the walk reaches them
  for (; w->visit_at < w->live_count; w->visit_at++) {
    if (w->visit_at + AHEAD_VISITS < w->live_count) __builtin_prefetch(&visits[w->live[w->visit_at + AHEAD_VISITS]]);
    const int i = w->live[w->visit_at];
    Visit& v = visits[i];
    Manifold& m = w->rebuilt[i];
    if (!v.again && (w->fresh[v.a] || w->fresh[v.b])) {
      v.again = true;
      prepare_visit(w, v, m);
      if (v.kind >= 2) {
        fill_rows(w);  // before the device runs the batches again
        return true;
      }
    }
    if ((rc = finish_visit(w, v, m)) != BM_OK) {
      w->fills.clear();
      return false;
    }
  }
  fill_rows(w);
  for (int s : w->woken) w->fresh[s] = 0;
  w->woken.clear();
  return false;
}


// src/world.js:1369-1382 (_syncPairs), with no bodies removed and no static body moved by hand.
void sync_pairs(bm_world* w) {
  Pairs& pr = w->pairs;
  while (w->wake_pending) {
    w->wake_pending = false;
    for (int p = pr.active; p < pr.count; p++) {
      if (w->awake[pr.a[p]] == 1 || w->awake[pr.b[p]] == 1) pr.wake(p);
    }
  }
}

// ---- freezing and sleeping ----

void join(std::vector<int32_t>& parent, const std::vector<uint8_t>& type, int a, int b);
int root_of(std::vector<int32_t>& parent, int s) {
  while (parent[s] != s) {
    parent[s] = parent[parent[s]];
    s = parent[s];
  }
  return s;
}
// src/world.js:2304-2309
void join(std::vector<int32_t>& parent, const std::vector<uint8_t>& type, int a, int b) {
  if (type[a] != DYNAMIC || type[b] != DYNAMIC) return;
  const int ra = root_of(parent, a), rb = root_of(parent, b);
  if (ra < rb) parent[rb] = ra;
  else if (rb < ra) parent[ra] = rb;
}

// src/world.js:1116-1172 (_freeze)
void freeze(bm_world* w) {
  const int n = w->awake_count, count = w->c.count;
  const int* A = w->c.a.data();
  const int* B = w->c.b.data();
  auto& calm = w->calm;
  auto& type = w->type;
  for (int i = 0; i < n; i++) {
    const int s = w->awake_list[i];
    calm[s] = w->still[s] >= SLEEP_TIME && (calm[s] & HANDLED) == 0 ? (STILL | QUIET | CALM) : 0;
  }
  for (int m = 0; m < count; m++) {
    const int a = A[m], b = B[m];
    if (type[b] != STATIC && (calm[b] & STILL) == 0) calm[a] &= ~(QUIET | CALM);
    if (type[a] != STATIC && (calm[a] & STILL) == 0) calm[b] &= ~(QUIET | CALM);
  }
  for (int m = 0; m < count; m++) {
    const int a = A[m], b = B[m];
    if (type[b] != STATIC && (calm[b] & QUIET) == 0) calm[a] &= ~CALM;
    if (type[a] != STATIC && (calm[a] & QUIET) == 0) calm[b] &= ~CALM;
  }
  if (w->frozen_count > 0) {
    auto& parent = w->parent;
    for (int i = 0; i < n; i++) parent[w->awake_list[i]] = w->awake_list[i];
    for (int m = 0; m < count; m++)
      if (w->frozen[A[m]] != 0 && w->frozen[B[m]] != 0) join(parent, type, A[m], B[m]);
    bool thawing = false;
    for (int i = 0; i < n; i++) {
      const int s = w->awake_list[i];
      if (w->frozen[s] != 0 && (calm[s] & CALM) == 0) { calm[root_of(parent, s)] |= THAWING; thawing = true; }
    }
    if (thawing) {
      for (int i = 0; i < n; i++) {
        const int s = w->awake_list[i];
        if (w->frozen[s] != 0 && (calm[root_of(parent, s)] & THAWING) != 0) calm[s] &= ~CALM;
      }
    }
  }
  for (int i = 0; i < n; i++) {
    const int s = w->awake_list[i];
    const bool fr = (calm[s] & CALM) != 0 && type[s] == DYNAMIC && w->joint_head[s] == 0;
    if (fr && w->frozen[s] == 0) {
      w->frozen[s] = 2;
      w->frozen_count++;
      w->inv_mass[s] = 0;
      for (int k = 0; k < 3; k++) { w->velocity[3 * s + k] = 0; w->angular[3 * s + k] = 0; }
    } else if (!fr && w->frozen[s] != 0) {
      thaw(w, s, true);
    }
  }
}

// src/world.js:1271-1286 (_fallAsleep)
void fall_asleep(bm_world* w, int s, int r) {
  if (w->frozen[s] != 0) thaw(w, s, false);
  for (int k = 0; k < 3; k++) { w->velocity[3 * s + k] = 0; w->angular[3 * s + k] = 0; }
  unawaken(w, s);
  w->island[s] = r + 1;
  if (s != r) {
    w->island_next[s] = w->island_next[r];
    w->island_next[r] = s + 1;
  }
  bounds_of(w, s, 0, 0);
  broadphase_settle(w, s);
  w->slept[w->slept_count++] = s;
}

// src/world.js:1196-1268 (_sleep), for worlds with no kinematic bodies, characters or joints.
void sleep_step(bm_world* w, double dt) {
  const double* p = w->position.data();
  const double* q = w->rotation.data();
  // Each awake body's timer and rest pose are its own: a large world's on the pool.
  auto timers = [&](int i0, int i1) {
  bool ready = false;
  for (int i = i0; i < i1; i++) {
    const int s = w->awake_list[i], s3 = 3 * s, s4 = 4 * s, r7 = 7 * s;
    const double* rest = &w->rest[r7];
    const double dx = p[s3] - rest[0], dy = p[s3 + 1] - rest[1], dz = p[s3 + 2] - rest[2];
    const double qx = q[s4], qy = q[s4 + 1], qz = q[s4 + 2], qw = q[s4 + 3];
    const double ax = rest[3], ay = rest[4], az = rest[5], aw = rest[6];
    const double tx = aw * qx - qw * ax - (qy * az - qz * ay);
    const double ty = aw * qy - qw * ay - (qz * ax - qx * az);
    const double tz = aw * qz - qw * az - (qx * ay - qy * ax);
    const double strayed = std::sqrt(dx * dx + dy * dy + dz * dz) + 2 * w->reach[s] * std::sqrt(tx * tx + ty * ty + tz * tz);
    if (strayed < w->size[s] * PIXEL) {
      const double before = w->still[s];
      w->still[s] += dt;
      if (w->still[s] >= SLEEP_TIME) ready = true;
      if (std::floor(w->still[s] / SLEEP_TIME) != std::floor(before / SLEEP_TIME)) {
        for (int k = 0; k < 3; k++) w->rest[r7 + k] = p[s3 + k];
        for (int k = 0; k < 4; k++) w->rest[r7 + 3 + k] = q[s4 + k];
      }
    } else {
      settle(w, s);
    }
  }
  return ready;
  };
  bool ready = false;
  const int n_awake = w->awake_count;
  if (w->pool && n_awake >= w->parallel_from) {
    const int chunk = 2048, chunks = (n_awake + chunk - 1) / chunk;
    std::vector<uint8_t> any((size_t)chunks, 0);
    w->pool->run(chunks, [&](int c) { any[c] = timers(c * chunk, std::min(n_awake, (c + 1) * chunk)) ? 1 : 0; });
    for (uint8_t a : any) ready = ready || a;
  } else {
    ready = timers(0, n_awake);
  }
  w->still_ready = ready;
  if (!ready) return;
  auto& parent = w->parent;
  const int n = w->awake_count;
  for (int i = 0; i < n; i++) parent[w->awake_list[i]] = w->awake_list[i];
  for (int m = 0; m < w->c.count; m++) join(parent, w->type, w->c.a[m], w->c.b[m]);
  for (int i = 0; i < n; i++) w->least[w->awake_list[i]] = INFINITY;
  for (int i = 0; i < n; i++) {
    const int s = w->awake_list[i], r = root_of(parent, s);
    if (w->still[s] < w->least[r]) w->least[r] = w->still[s];
  }
  bool slept = false;
  for (int i = n - 1; i >= 0; i--) {
    const int s = w->awake_list[i], r = root_of(parent, s);
    if (w->least[r] < SLEEP_TIME) continue;
    fall_asleep(w, s, r);
    slept = true;
  }
  if (!slept) return;
  Pairs& pr = w->pairs;
  for (int pp = 0; pp < pr.active;) {
    if (w->awake[pr.a[pp]] == 0 && w->awake[pr.b[pp]] == 0) pr.sleep(pp);
    else pp++;
  }
}

// ---- the continuous catch: src/world.js:1683-1743 (_catch), src/query.js:355-445 and
// 607-642 (advance, sweep), src/gjk.js (closest), src/contact.js:1755-1776 (gapOf) ----

constexpr int CORE_SPHERE = 0, CORE_CAPSULE = 1, CORE_BOX = 2, CORE_HULL = 3, CORE_TRIANGLE = 4;  // src/collide.js:44-49

// A shape as the sweep places it: kind, centre, axes (x, y, z columns), half extents; a
// hull's number, corners (V of them, in its own frame) and thinnest extent.
struct Core {
  int kind;
  double center[3], axes[9], half[3], weight[3];
  int hull = -1, V = 0;
  const double* vertices = nullptr;
  double hull_size = 0;
  double points[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // a triangle's corners, in the world
};

double core_radius(const Core& s) { return s.kind == CORE_SPHERE || s.kind == CORE_CAPSULE ? s.half[0] : 0; }

// src/gjk.js:70-107 (support) for a sphere's, a triangle's, a capsule's, a box's or a hull's core.
void core_support(const Core& s, double x, double y, double z, double* out) {
  const double* c = s.center;
  const double* a = s.axes;
  const double* h = s.half;
  if (s.kind == CORE_SPHERE) {
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
  } else if (s.kind == CORE_TRIANGLE) {
    // The corner farthest along; the first of equals.
    const double* p = s.points;
    int best = 0;
    double most = -INFINITY;
    for (int i = 0; i < 3; i++) {
      const double d = p[3 * i] * x + p[3 * i + 1] * y + p[3 * i + 2] * z;
      if (d > most) { most = d; best = i; }
    }
    out[0] = p[3 * best]; out[1] = p[3 * best + 1]; out[2] = p[3 * best + 2];
  } else if (s.kind == CORE_HULL) {
    // The corner farthest along, in the hull's own frame; the first of equals.
    const double* v = s.vertices;
    const double lx = x * a[0] + y * a[1] + z * a[2], ly = x * a[3] + y * a[4] + z * a[5], lz = x * a[6] + y * a[7] + z * a[8];
    int best = 0;
    double most = -INFINITY;
    for (int i = 0; i < s.V; i++) {
      const double d = v[3 * i] * lx + v[3 * i + 1] * ly + v[3 * i + 2] * lz;
      if (d > most) { most = d; best = i; }
    }
    const double px = v[3 * best], py = v[3 * best + 1], pz = v[3 * best + 2];
    out[0] = c[0] + px * a[0] + py * a[3] + pz * a[6];
    out[1] = c[1] + px * a[1] + py * a[4] + pz * a[7];
    out[2] = c[2] + px * a[2] + py * a[5] + pz * a[8];
  } else if (s.kind == CORE_CAPSULE) {
    const double e = x * a[3] + y * a[4] + z * a[5] >= 0 ? h[1] : -h[1];
    out[0] = c[0] + e * a[3]; out[1] = c[1] + e * a[4]; out[2] = c[2] + e * a[5];
  } else {
    const double ex = x * a[0] + y * a[1] + z * a[2] >= 0 ? h[0] : -h[0];
    const double ey = x * a[3] + y * a[4] + z * a[5] >= 0 ? h[1] : -h[1];
    const double ez = x * a[6] + y * a[7] + z * a[8] >= 0 ? h[2] : -h[2];
    out[0] = c[0] + ex * a[0] + ey * a[3] + ez * a[6];
    out[1] = c[1] + ex * a[1] + ey * a[4] + ez * a[7];
    out[2] = c[2] + ex * a[2] + ey * a[5] + ez * a[8];
  }
}

// GJK's simplex and its reduction to the feature nearest the origin (src/gjk.js:172-277).
struct Simplex {
  double W[12], PA[12], PB[12], L[4];
  int count;
  void move(int to, int from) {
    if (to == from) return;
    for (int c = 0; c < 3; c++) {
      W[3 * to + c] = W[3 * from + c]; PA[3 * to + c] = PA[3 * from + c]; PB[3 * to + c] = PB[3 * from + c];
    }
  }
  void keep(int i, int j, int k, double li, double lj, double lk, int n) {
    move(0, i); L[0] = li;
    if (n > 1) { move(1, j); L[1] = lj; }
    if (n > 2) { move(2, k); L[2] = lk; }
    count = n;
  }
  void segment(int i, int j) {
    const double ax = W[3 * i], ay = W[3 * i + 1], az = W[3 * i + 2];
    const double ex = W[3 * j] - ax, ey = W[3 * j + 1] - ay, ez = W[3 * j + 2] - az;
    const double ee = ex * ex + ey * ey + ez * ez;
    const double t = ee > 0 ? -(ax * ex + ay * ey + az * ez) / ee : 0;
    if (t <= 0) keep(i, i, i, 1, 0, 0, 1);
    else if (t >= 1) keep(j, j, j, 1, 0, 0, 1);
    else keep(i, j, j, 1 - t, t, 0, 2);
  }
  void triangle(int i, int j, int k) {
    const double ax = W[3 * i], ay = W[3 * i + 1], az = W[3 * i + 2];
    const double bx = W[3 * j], by = W[3 * j + 1], bz = W[3 * j + 2];
    const double cx = W[3 * k], cy = W[3 * k + 1], cz = W[3 * k + 2];
    const double abx = bx - ax, aby = by - ay, abz = bz - az, acx = cx - ax, acy = cy - ay, acz = cz - az;
    const double d1 = -(abx * ax + aby * ay + abz * az), d2 = -(acx * ax + acy * ay + acz * az);
    if (d1 <= 0 && d2 <= 0) return keep(i, i, i, 1, 0, 0, 1);
    const double d3 = -(abx * bx + aby * by + abz * bz), d4 = -(acx * bx + acy * by + acz * bz);
    if (d3 >= 0 && d4 <= d3) return keep(j, j, j, 1, 0, 0, 1);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      const double v = d1 / (d1 - d3);
      return keep(i, j, j, 1 - v, v, 0, 2);
    }
    const double d5 = -(abx * cx + aby * cy + abz * cz), d6 = -(acx * cx + acy * cy + acz * cz);
    if (d6 >= 0 && d5 <= d6) return keep(k, k, k, 1, 0, 0, 1);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
      const double w = d2 / (d2 - d6);
      return keep(i, k, k, 1 - w, w, 0, 2);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
      const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return keep(j, k, k, 1 - w, w, 0, 2);
    }
    const double sc = 1 / (va + vb + vc), v = vb * sc, w = vc * sc;
    keep(i, j, k, 1 - v - w, v, w, 3);
  }
  bool tetrahedron() {
    static const uint8_t FACES[16] = {0, 1, 2, 3, 0, 1, 3, 2, 0, 2, 3, 1, 1, 2, 3, 0};  // src/gjk.js:60
    const double FLAT = 1.0 / 4503599627370496;                                       // 2^-52
    double best = INFINITY;
    int best_face = -1;
    double w[12], pa[12], pb[12];
    std::memcpy(w, W, sizeof w); std::memcpy(pa, PA, sizeof pa); std::memcpy(pb, PB, sizeof pb);
    bool outside = false;
    for (int f = 0; f < 4; f++) {
      const int i = FACES[4 * f], j = FACES[4 * f + 1], k = FACES[4 * f + 2], o = FACES[4 * f + 3];
      const double ax = w[3 * i], ay = w[3 * i + 1], az = w[3 * i + 2];
      const double ux = w[3 * j] - ax, uy = w[3 * j + 1] - ay, uz = w[3 * j + 2] - az;
      const double vx = w[3 * k] - ax, vy = w[3 * k + 1] - ay, vz = w[3 * k + 2] - az;
      const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      const double to_origin = -(nx * ax + ny * ay + nz * az);
      const double ox = w[3 * o] - ax, oy = w[3 * o + 1] - ay, oz = w[3 * o + 2] - az;
      const double to_fourth = nx * ox + ny * oy + nz * oz;
      const bool flat = to_fourth * to_fourth <= FLAT * (nx * nx + ny * ny + nz * nz) * (ox * ox + oy * oy + oz * oz);
      if (!flat && to_origin * to_fourth >= 0) continue;
      outside = true;
      std::memcpy(W, w, sizeof w); std::memcpy(PA, pa, sizeof pa); std::memcpy(PB, pb, sizeof pb);
      count = 4;
      triangle(i, j, k);
      double px = 0, py = 0, pz = 0;
      for (int m = 0; m < count; m++) {
        px += L[m] * W[3 * m]; py += L[m] * W[3 * m + 1]; pz += L[m] * W[3 * m + 2];
      }
      const double d = px * px + py * py + pz * pz;
      if (d < best) { best = d; best_face = f; }
    }
    if (!outside) return false;
    const int i = FACES[4 * best_face], j = FACES[4 * best_face + 1], k = FACES[4 * best_face + 2];
    std::memcpy(W, w, sizeof w); std::memcpy(PA, pa, sizeof pa); std::memcpy(PB, pb, sizeof pb);
    count = 4;
    triangle(i, j, k);
    return true;
  }
  bool reduce() {
    if (count == 2) { segment(0, 1); return true; }
    if (count == 3) { triangle(0, 1, 2); return true; }
    return tetrahedron();
  }
};

// src/gjk.js:114-155 (closest)
double closest(const Core& a, const Core& b, double* near_a, double* near_b) {
  const double PROGRESS = 1.0 / 1073741824;  // 2^-30
  Simplex S;
  double sa[3], sb[3];
  double vx = a.center[0] - b.center[0], vy = a.center[1] - b.center[1], vz = a.center[2] - b.center[2];
  if (vx == 0 && vy == 0 && vz == 0) vx = 1;
  core_support(a, -vx, -vy, -vz, sa);
  core_support(b, vx, vy, vz, sb);
  for (int c = 0; c < 3; c++) { S.W[c] = sa[c] - sb[c]; S.PA[c] = sa[c]; S.PB[c] = sb[c]; }
  S.count = 1;
  S.L[0] = 1;
  vx = S.W[0]; vy = S.W[1]; vz = S.W[2];
  for (int round = 0; round < 64; round++) {
    const double vv = vx * vx + vy * vy + vz * vz;
    if (vv == 0) break;
    core_support(a, -vx, -vy, -vz, sa);
    core_support(b, vx, vy, vz, sb);
    const double wx = sa[0] - sb[0], wy = sa[1] - sb[1], wz = sa[2] - sb[2];
    if (vv - (vx * wx + vy * wy + vz * wz) <= PROGRESS * vv) break;
    bool repeat = false;
    for (int i = 0; i < S.count; i++)
      if (S.W[3 * i] == wx && S.W[3 * i + 1] == wy && S.W[3 * i + 2] == wz) repeat = true;
    if (repeat) break;
    const int i3 = 3 * S.count;
    S.W[i3] = wx; S.W[i3 + 1] = wy; S.W[i3 + 2] = wz;
    for (int c = 0; c < 3; c++) { S.PA[i3 + c] = sa[c]; S.PB[i3 + c] = sb[c]; }
    S.count++;
    if (!S.reduce()) { vx = vy = vz = 0; break; }
    vx = 0; vy = 0; vz = 0;
    for (int i = 0; i < S.count; i++) {
      vx += S.L[i] * S.W[3 * i]; vy += S.L[i] * S.W[3 * i + 1]; vz += S.L[i] * S.W[3 * i + 2];
    }
  }
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  for (int i = 0; i < S.count; i++) {
    ax += S.L[i] * S.PA[3 * i]; ay += S.L[i] * S.PA[3 * i + 1]; az += S.L[i] * S.PA[3 * i + 2];
    bx += S.L[i] * S.PB[3 * i]; by += S.L[i] * S.PB[3 * i + 1]; bz += S.L[i] * S.PB[3 * i + 2];
  }
  near_a[0] = ax; near_a[1] = ay; near_a[2] = az;
  near_b[0] = bx; near_b[1] = by; near_b[2] = bz;
  return std::sqrt(vx * vx + vy * vy + vz * vz);
}

// src/collide.js:82-91 (Shape.place)
void place_at(Core& s, double x, double y, double z, const double* q) {
  s.center[0] = x; s.center[1] = y; s.center[2] = z;
  axes_of(q, s.axes);
}

// src/query.js:91-95 (place): slot s's shape where it is now.
void place_body(const bm_world* w, int s, Core& out) {
  const int s3 = 3 * s;
  out.hull = -1;
  if (w->shape[s] == HULL) {
    const auto& h = w->hulls[w->hull[s]];
    out.kind = CORE_HULL;
    out.hull = w->hull[s];
    out.V = h.V;
    out.vertices = &w->hull_vertices[3 * (size_t)h.v0];
    out.hull_size = h.size;
    out.half[0] = out.half[1] = out.half[2] = 0;
  } else if (w->shape[s] == BOX) {
    out.kind = CORE_BOX;
    for (int k = 0; k < 3; k++) out.half[k] = w->dims[s3 + k] * 0.5;
  } else if (w->shape[s] == CAPSULE) {
    out.kind = CORE_CAPSULE;
    out.half[0] = w->dims[s3]; out.half[1] = (w->dims[s3 + 1] - 2 * w->dims[s3]) * 0.5; out.half[2] = 0;
  } else {
    out.kind = CORE_SPHERE;
    out.half[0] = w->dims[s3]; out.half[1] = 0; out.half[2] = 0;
  }
  const double scale = w->type[s] == DYNAMIC ? w->gravity_scale[s] : 0;
  for (int k = 0; k < 3; k++) out.weight[k] = w->gravity[k] * scale;
  place_at(out, w->position[s3], w->position[s3 + 1], w->position[s3 + 2], &w->rotation[4 * s]);
}

// src/collide.js:168-177 (reachAlong) for a hull: how far its corners reach along unit (x, y, z).
double hull_reach(const Core& s, double x, double y, double z) {
  const double* A = s.axes;
  const double* v = s.vertices;
  const double lx = A[0] * x + A[1] * y + A[2] * z, ly = A[3] * x + A[4] * y + A[5] * z, lz = A[6] * x + A[7] * y + A[8] * z;
  double most = -INFINITY;
  for (int i = 0; i < s.V; i++) most = js_max(most, v[3 * i] * lx + v[3 * i + 1] * ly + v[3 * i + 2] * lz);
  return most;
}

// src/query.js:114-128 (extents) for a sphere, a capsule, a box or a hull.
void extents(const Core& s, double* out) {
  const double* a = s.axes;
  const double* h = s.half;
  if (s.kind == CORE_SPHERE) {
    out[0] = out[1] = out[2] = h[0];
  } else if (s.kind == CORE_CAPSULE) {
    for (int k = 0; k < 3; k++) out[k] = std::fabs(a[3 + k]) * h[1] + h[0];
  } else if (s.kind == CORE_HULL) {
    for (int k = 0; k < 3; k++)
      out[k] = js_max(hull_reach(s, k == 0 ? 1 : 0, k == 1 ? 1 : 0, k == 2 ? 1 : 0),
                      hull_reach(s, k == 0 ? -1 : 0, k == 1 ? -1 : 0, k == 2 ? -1 : 0));
  } else {
    for (int k = 0; k < 3; k++)
      out[k] = std::fabs(a[k]) * h[0] + std::fabs(a[3 + k]) * h[1] + std::fabs(a[6 + k]) * h[2];
  }
}

// src/query.js:138-167 (strays), listed once a step.
const std::vector<int32_t>& strays(bm_world* w) {
  if (w->strays_at == w->step_count) return w->strays;
  w->strays.clear();
  const double* fat = w->trees[MOVING_TREE].box.data();
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (w->type[s] != DYNAMIC) continue;
    const int o = 6 * w->leaf[s];
    Core body;
    place_body(w, s, body);
    double ext[3];
    extents(body, ext);
    const double* c = body.center;
    if (c[0] - ext[0] < fat[o] || c[1] - ext[1] < fat[o + 1] || c[2] - ext[2] < fat[o + 2] || c[0] + ext[0] > fat[o + 3] ||
        c[1] + ext[1] > fat[o + 4] || c[2] + ext[2] > fat[o + 5])
      w->strays.push_back(s);
  }
  w->strays_at = w->step_count;
  return w->strays;
}

double core_size(const Core& s) {
  if (s.kind == CORE_SPHERE || s.kind == CORE_CAPSULE) return 2 * s.half[0];
  if (s.kind == CORE_HULL) return s.hull_size;
  return 2 * js_min(js_min(s.half[0], s.half[1]), s.half[2]);
}

// src/triangle.js:104-124 and 240-334 (touch, nearest, sphere) for a ball against a triangle
// met alone, as the catch's sweep meets one (src/query.js:37 BARE: every edge a corner from
// both sides, no folds, two-sided), with no margin: the normal from the ball to the
// triangle and the contact point; false when apart.
bool touch_ball(const Core& s, const Core& t, double tolerance, double* normal, double* point) {
  const double* P = t.points;
  const double nx0 = t.axes[0], ny0 = t.axes[1], nz0 = t.axes[2];
  const double* c = s.center;
  const double r = s.half[0];
  const bool front = (c[0] - P[0]) * nx0 + (c[1] - P[1]) * ny0 + (c[2] - P[2]) * nz0 >= 0;
  const double k = front ? 1 : -1;
  const double up[3] = {k * nx0, k * ny0, k * nz0};
  // Ericson 5.1.5, the closest point on the triangle (src/triangle.js:240-270).
  double Q[3];
  {
    const double ax = P[0], ay = P[1], az = P[2];
    const double abx = P[3] - ax, aby = P[4] - ay, abz = P[5] - az, acx = P[6] - ax, acy = P[7] - ay, acz = P[8] - az;
    const double apx = c[0] - ax, apy = c[1] - ay, apz = c[2] - az;
    const double d1 = abx * apx + aby * apy + abz * apz, d2 = acx * apx + acy * apy + acz * apz;
    double u, v;
    const double bpx = c[0] - P[3], bpy = c[1] - P[4], bpz = c[2] - P[5];
    const double d3 = abx * bpx + aby * bpy + abz * bpz, d4 = acx * bpx + acy * bpy + acz * bpz;
    const double vc = d1 * d4 - d3 * d2;
    const double cpx = c[0] - P[6], cpy = c[1] - P[7], cpz = c[2] - P[8];
    const double d5 = abx * cpx + aby * cpy + abz * cpz, d6 = acx * cpx + acy * cpy + acz * cpz;
    const double vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    if (d1 <= 0 && d2 <= 0) { u = 0; v = 0; }
    else if (d3 >= 0 && d4 <= d3) { u = 1; v = 0; }
    else if (vc <= 0 && d1 >= 0 && d3 <= 0) { u = d1 / (d1 - d3); v = 0; }
    else if (d6 >= 0 && d5 <= d6) { u = 0; v = 1; }
    else if (vb <= 0 && d2 >= 0 && d6 <= 0) { u = 0; v = d2 / (d2 - d6); }
    else if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
      const double wt = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      u = 1 - wt; v = wt;
    } else {
      const double den = 1 / (va + vb + vc);
      u = vb * den; v = vc * den;
    }
    Q[0] = ax + u * abx + v * acx; Q[1] = ay + u * aby + v * acy; Q[2] = az + u * abz + v * acz;
  }
  double nx = c[0] - Q[0], ny = c[1] - Q[1], nz = c[2] - Q[2];
  const double d = std::sqrt(nx * nx + ny * ny + nz * nz);
  double sep;
  if (d > r / 67108864) {
    nx /= d; ny /= d; nz /= d;
    sep = d - r;
  } else {
    nx = up[0]; ny = up[1]; nz = up[2];
    sep = (c[0] - P[0]) * up[0] + (c[1] - P[1]) * up[1] + (c[2] - P[2]) * up[2] - r;
  }
  if (sep > 0) return false;
  normal[0] = -nx; normal[1] = -ny; normal[2] = -nz;
  const double m = r + sep / 2;
  point[0] = c[0] - nx * m; point[1] = c[1] - ny * m; point[2] = c[2] - nz * m;
  return true;
}

// src/mesh.js:317-331 (grid, triangle): triangle t of field F's corners, in its body's frame.
void field_corners(const bm_world* w, const bm_world::HostField& F, int t, double* out) {
  const int cell = t >> 1, cells = F.columns - 1, i = cell % cells, j = (cell - i) / cells;
  int g[6];
  if ((t & 1) == 0) { g[0] = i; g[1] = j; g[2] = i + 1; g[3] = j + 1; g[4] = i + 1; g[5] = j; }
  else { g[0] = i; g[1] = j; g[2] = i; g[3] = j + 1; g[4] = i + 1; g[5] = j + 1; }
  const double* h = w->field_heights.data();
  for (int c = 0; c < 3; c++) {
    out[3 * c] = F.x0 + g[2 * c] * F.spacing;
    out[3 * c + 1] = F.at(h, g[2 * c], g[2 * c + 1]);
    out[3 * c + 2] = F.z0 + g[2 * c + 1] * F.spacing;
  }
}

// src/query.js:56-66 (triangle): triangle t of slot s's field placed in the world.
void field_triangle(const bm_world* w, int s, int t, Core& tri) {
  const auto& F = w->fields[w->field[s]];
  double* p = tri.points;
  field_corners(w, F, t, p);
  const double* q = &w->rotation[4 * s];
  const double* pos = &w->position[3 * s];
  for (int v = 0; v < 9; v += 3) {
    double l[3];
    rotate(q, p[v], p[v + 1], p[v + 2], false, l);
    p[v] = pos[0] + l[0]; p[v + 1] = pos[1] + l[1]; p[v + 2] = pos[2] + l[2];
  }
  for (int k = 0; k < 3; k++) tri.center[k] = (p[k] + p[k + 3] + p[k + 6]) / 3;
  // src/mesh.js:104-113 (normal): left as it was for a triangle with no area.
  const double ux = p[3] - p[0], uy = p[4] - p[1], uz = p[5] - p[2];
  const double vx = p[6] - p[0], vy = p[7] - p[1], vz = p[8] - p[2];
  const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (l > 0) { tri.axes[0] = nx / l; tri.axes[1] = ny / l; tri.axes[2] = nz / l; }
}

// src/mesh.js:388-401 (triangles): both triangles of every cell under the box whose heights reach it.
template <class V>
void field_triangles(const bm_world* w, const bm_world::HostField& F, double x0, double y0, double z0, double x1,
                     double y1, double z1, V&& visit) {
  const double s = F.spacing;
  const int cells = F.columns - 1;
  const double i0 = js_max(0, std::floor((x0 - F.x0) / s)), i1 = js_min(cells - 1, std::floor((x1 - F.x0) / s));
  const double j0 = js_max(0, std::floor((z0 - F.z0) / s)), j1 = js_min(F.rows - 2, std::floor((z1 - F.z0) / s));
  if (!(i0 <= i1) || !(j0 <= j1)) return;
  const double* h = w->field_heights.data();
  for (int j = (int)j0; j <= (int)j1; j++) {
    for (int i = (int)i0; i <= (int)i1; i++) {
      const double a = F.at(h, i, j), b = F.at(h, i + 1, j), c = F.at(h, i, j + 1), d = F.at(h, i + 1, j + 1);
      if (js_max(js_max(js_max(a, b), c), d) < y0 || js_min(js_min(js_min(a, b), c), d) > y1) continue;
      const int t = 2 * (j * cells + i);
      visit(t);
      visit(t + 1);
    }
  }
}

// What advance() found: where along the sweep, the normal (body towards cast) and the point.
struct Hit {
  double normal[3], point[3];
};

// src/query.js:361-406 (advance): conservative advancement of `cast` from o along d against
// b, up to max_t; the distance, or -1. A core already meeting b asks the device narrowphase
// which way out, as the reference asks collide(). rc is set when that call fails.
double advance(bm_world* w, Core& cast, const Core& b, double scale, const double* o, const double* d, double max_t,
               bool sliding, Hit& hit, int& rc) {
  const double ra = core_radius(cast), rb = core_radius(b), tolerance = scale * PIXEL;
  const double meeting = scale / 67108864;  // 2^-26
  double t = 0;
  for (int i = 0; i < 64; i++) {
    cast.center[0] = o[0] + t * d[0]; cast.center[1] = o[1] + t * d[1]; cast.center[2] = o[2] + t * d[2];
    double na[3], nb[3];
    const double dist = closest(cast, b, na, nb), gap = dist - ra - rb;
    if (gap < tolerance) {
      if (sliding && i == 0 && dist > meeting) {
        const double opening = ((na[0] - nb[0]) * d[0] + (na[1] - nb[1]) * d[1] + (na[2] - nb[2]) * d[2]) / dist;
        if (opening >= 0 || gap + opening * max_t >= js_min(gap, 0) - meeting) return -1;
      }
      if (dist > meeting) {
        const double nx = (na[0] - nb[0]) / dist, ny = (na[1] - nb[1]) / dist, nz = (na[2] - nb[2]) / dist;
        hit.normal[0] = nx; hit.normal[1] = ny; hit.normal[2] = nz;
        hit.point[0] = nb[0] + nx * rb; hit.point[1] = nb[1] + ny * rb; hit.point[2] = nb[2] + nz * rb;
        return t;
      }
      if (b.kind == CORE_TRIANGLE) {
        // src/query.js:388: a triangle met alone (seams BARE, no folds) asks touch().
        if (cast.kind != CORE_SPHERE)
          return rc = fail(w, BM_ERR_UNSUPPORTED, "step: a swept shape other than a ball deep in level geometry"), -1;
        double n[3], at[3];
        if (touch_ball(cast, b, tolerance, n, at)) {
          for (int k = 0; k < 3; k++) { hit.normal[k] = -n[k]; hit.point[k] = at[k]; }
        } else {
          for (int k = 0; k < 3; k++) { hit.normal[k] = -d[k]; hit.point[k] = cast.center[k]; }
        }
        return t;
      }
      double ra_rec[19], rb_rec[19];
      const Core* both[2] = {&cast, &b};
      double* recs[2] = {ra_rec, rb_rec};
      for (int k = 0; k < 2; k++) {
        const Core& c = *both[k];
        double* r = recs[k];
        r[0] = c.kind;
        for (int j = 0; j < 3; j++) r[1 + j] = c.center[j];
        for (int j = 0; j < 9; j++) r[4 + j] = c.axes[j];
        for (int j = 0; j < 3; j++) r[13 + j] = c.half[j];
        if (c.kind == CORE_HULL) r[13] = w->hull_base + c.hull;
        for (int j = 0; j < 3; j++) r[16 + j] = c.weight[j];
      }
      const double zero = 0;
      int32_t count = 0;
      double normal[3], points[12], seps[4];
      uint32_t ids[4];
      {
        // A batch's worlds share its device context: one call at a time.
        std::unique_lock<std::mutex> held;
        if (w->device_lock) held = std::unique_lock<std::mutex>(*w->device_lock);
        rc = bm_collide_host(w->ctx, 1, ra_rec, rb_rec, &zero, &tolerance, &count, normal, points, seps, ids, nullptr);
      }
      if (rc != BM_OK) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx)), -1;
      if (count > 0) {
        for (int k = 0; k < 3; k++) { hit.normal[k] = -normal[k]; hit.point[k] = points[k]; }
      } else {
        for (int k = 0; k < 3; k++) { hit.normal[k] = -d[k]; hit.point[k] = cast.center[k]; }
      }
      return t;
    }
    const double closing = -((na[0] - nb[0]) * d[0] + (na[1] - nb[1]) * d[1] + (na[2] - nb[2]) * d[2]) / dist;
    if (closing <= 0) return -1;
    t += gap / closing;
    if (t >= max_t) return -1;
  }
  return -1;
}

// src/query.js:607-642 (sweep, sweepVisit) for a world of plain bodies: the first slot
// the cast meets sliding from its centre along unit d for up to max_t, or -1; out gets
// the distance, the normal and the point. Slots `skip` marks are passed over.
// src/query.js:644-668 (sweepLevel, sweepTriangle) and src/mesh.js:341-363 (Heightfield.ray,
// a cast of any size): `cast` slid from o along d against the triangles of slot s's field
// under the box the cast sweeps through, each by advance(); the nearest touch's distance, or -1.
double sweep_level(bm_world* w, Core& cast, int s, const double* o, const double* d, const double* g, double max_t,
                   Hit& hit, int& rc) {
  const double* q = &w->rotation[4 * s];
  const double* pos = &w->position[3 * s];
  double l[3], u[3], half[3] = {0, 0, 0};
  rotate(q, o[0] - pos[0], o[1] - pos[1], o[2] - pos[2], true, l);
  rotate(q, d[0], d[1], d[2], true, u);
  // src/query.js:73-81 (extentsIn)
  for (int j = 0; j < 3; j++) {
    double e[3];
    rotate(q, j == 0 ? 1 : 0, j == 1 ? 1 : 0, j == 2 ? 1 : 0, true, e);
    for (int k = 0; k < 3; k++) half[k] += std::fabs(e[k]) * g[j];
  }
  const auto& F = w->fields[w->field[s]];
  double t0 = 0, t1 = max_t;
  for (int k = 0; k < 3; k++) {
    if (u[k] == 0) {
      if (l[k] < F.lo[k] - half[k] || l[k] > F.hi[k] + half[k]) return -1;
      continue;
    }
    double near = (F.lo[k] - half[k] - l[k]) / u[k], far = (F.hi[k] + half[k] - l[k]) / u[k];
    if (near > far) { const double x = near; near = far; far = x; }
    if (near > t0) t0 = near;
    if (far < t1) t1 = far;
    if (t0 > t1) return -1;
  }
  const double ax = l[0] + t0 * u[0], ay = l[1] + t0 * u[1], az = l[2] + t0 * u[2];
  const double bx = l[0] + t1 * u[0], by = l[1] + t1 * u[1], bz = l[2] + t1 * u[2];
  int best = -1;
  double level_max = max_t;
  Hit level_hit{};
  const double scale = core_size(cast);
  field_triangles(w, F, js_min(ax, bx) - half[0], js_min(ay, by) - half[1], js_min(az, bz) - half[2],
                  js_max(ax, bx) + half[0], js_max(ay, by) + half[1], js_max(az, bz) + half[2], [&](int t) {
                    if (rc != BM_OK) return;
                    Core tri;
                    tri.kind = CORE_TRIANGLE;
                    for (int k = 0; k < 9; k++) tri.axes[k] = (k == 0 || k == 4 || k == 8) ? 1 : 0;
                    for (int k = 0; k < 3; k++) { tri.half[k] = 0; tri.weight[k] = 0; }
                    field_triangle(w, s, t, tri);
                    Hit h;
                    const double at = advance(w, cast, tri, scale, o, d, level_max, true, h, rc);
                    if (at >= 0) { best = t; level_max = at; level_hit = h; }
                  });
  if (best < 0) return -1;
  hit = level_hit;
  return level_max;
}

int sweep(bm_world* w, Core& cast, int self, const double* d, double max_t, double* out, const uint8_t* skip,
          int skipped, int& rc, std::vector<int>* stack = nullptr, std::vector<double>* at = nullptr) {
  const double o[3] = {cast.center[0], cast.center[1], cast.center[2]};
  double g[3];
  extents(cast, g);
  int best = -1;
  Hit best_hit{};
  auto visit = [&](int s) -> double {
    if (rc != BM_OK || skip[s] != 0 || s == self) return max_t;
    Hit hit;
    double t;
    if (w->shape[s] >= MESH) {
      t = sweep_level(w, cast, s, o, d, g, max_t, hit, rc);
    } else {
      Core body;
      place_body(w, s, body);
      const double own = w->size[s], scale = own > 0 ? js_min(own, core_size(cast)) : core_size(cast);
      t = advance(w, cast, body, scale, o, d, max_t, true, hit, rc);
    }
    if (t < 0) return max_t;
    best = s;
    max_t = t;
    best_hit = hit;
    return max_t;
  };
  for (int k = 0; k < 2; k++) {
    if (k == MOVING_TREE && w->trees[k].leaves == skipped) continue;
    max_t = stack ? w->trees[k].raycast(o, d, max_t, g, visit, *stack, *at) : w->trees[k].raycast(o, d, max_t, g, visit);
  }
  for (int s : strays(w)) visit(s);
  cast.center[0] = o[0]; cast.center[1] = o[1]; cast.center[2] = o[2];
  if (best >= 0) {
    out[0] = max_t;
    for (int k = 0; k < 3; k++) { out[1 + k] = best_hit.normal[k]; out[4 + k] = best_hit.point[k]; }
  }
  return best;
}

// src/contact.js:1755-1776 (gapOf): manifold m's least separation where its bodies are now.
double gap_of(const bm_world* w, int m) {
  const Contacts& c = w->c;
  const double* p = w->position.data();
  const double* start = w->start.data();
  const int a = c.a[m], b = c.b[m], a3 = 3 * a, b3 = 3 * b, m3 = 3 * m;
  const double nx = c.normal[m3], ny = c.normal[m3 + 1], nz = c.normal[m3 + 2];
  const double dx = (p[b3] - start[b3]) - (p[a3] - start[a3]);
  const double dy = (p[b3 + 1] - start[b3 + 1]) - (p[a3 + 1] - start[a3 + 1]);
  const double dz = (p[b3 + 2] - start[b3 + 2]) - (p[a3 + 2] - start[a3 + 2]);
  const double dn = dx * nx + dy * ny + dz * nz;
  double na[3], nb[3];
  rotate(&w->rotation[4 * a], nx, ny, nz, true, na);
  rotate(&w->rotation[4 * b], nx, ny, nz, true, nb);
  double least_gap = INFINITY;
  for (int k = 0; k < c.points[m]; k++) {
    const int i = m * P + k, i3 = 3 * i;
    const double s = dn + (c.lb[i3] * nb[0] + c.lb[i3 + 1] * nb[1] + c.lb[i3 + 2] * nb[2])
        - (c.la[i3] * na[0] + c.la[i3 + 1] * na[1] + c.la[i3 + 2] * na[2]) + c.adjusted[i];
    if (s < least_gap) least_gap = s;
  }
  return least_gap;
}

// src/world.js:1683-1743 (_catch), for a world with no joints.
int catch_fast(bm_world* w) {
  const int n = w->awake_count;
  const int* list = w->awake_list.data();
  double* p = w->position.data();
  double* q = w->rotation.data();
  const double* st = w->start.data();
  const double* sq = w->start_q.data();
  int count = 0;
  for (int i = 0; i < n; i++) {
    const int s = list[i], s3 = 3 * s;
    if (w->type[s] != DYNAMIC) continue;
    const double dx = p[s3] - st[s3], dy = p[s3 + 1] - st[s3 + 1], dz = p[s3 + 2] - st[s3 + 2], half = 0.5 * w->size[s];
    if (dx * dx + dy * dy + dz * dz > half * half) { w->fast[s] = 1; count++; }
  }
  if (count == 0) return BM_OK;
  w->catch_steps++;
  // Each fast body sweeps against the bodies that are not fast (the fast ones are skipped), and
  // a catch moves only the body caught: the sweeps are independent, so a large world's run on
  // its host threads, in any order, with the same results. Device calls (a deep core's way
  // out) then take turns.
  std::vector<int32_t> fast_list;
  for (int i = 0; i < n; i++)
    if (w->fast[list[i]] == 1) fast_list.push_back(list[i]);
  std::vector<int> rcs(fast_list.size(), BM_OK);
  std::vector<uint8_t> hits(fast_list.size(), 0);
  std::mutex turns;
  std::mutex* lock = w->device_lock;
  const bool threads = w->pool && (int)fast_list.size() >= CATCH_PARALLEL;
  if (threads) {
    strays(w);  // listed once, before any body moves, as the first sweep would
    if (!w->device_lock) w->device_lock = &turns;
  }
  auto one = [&](int k) {
    const int s = fast_list[k], s3 = 3 * s, s4 = 4 * s;
    int& rc = rcs[k];
    const double dx = p[s3] - st[s3], dy = p[s3 + 1] - st[s3 + 1], dz = p[s3 + 2] - st[s3 + 2];
    const double length = std::sqrt(dx * dx + dy * dy + dz * dz);
    Core core{};
    core.kind = CORE_SPHERE;
    core.half[0] = 0.125 * w->size[s];
    const double identity[4] = {0, 0, 0, 1};
    place_at(core, st[s3], st[s3 + 1], st[s3 + 2], identity);
    const double dir[3] = {dx / length, dy / length, dz / length};
    double swept[7];
    thread_local std::vector<int> stack;
    thread_local std::vector<double> at;
    const int hit = sweep(w, core, s, dir, length, swept, w->fast.data(), count, rc, &stack, &at);
    if (rc != BM_OK || hit < 0) return;
    const double f = swept[0] / length;
    p[s3] = st[s3] + dx * f; p[s3 + 1] = st[s3 + 1] + dy * f; p[s3 + 2] = st[s3 + 2] + dz * f;
    const double sign = sq[s4] * q[s4] + sq[s4 + 1] * q[s4 + 1] + sq[s4 + 2] * q[s4 + 2] + sq[s4 + 3] * q[s4 + 3] < 0 ? -1 : 1;
    double turned[4], l = 0;
    for (int k = 0; k < 4; k++) {
      const double v = (1 - f) * sq[s4 + k] + f * sign * q[s4 + k];
      turned[k] = v;
      l += v * v;
    }
    l = 1 / std::sqrt(l);
    for (int k = 0; k < 4; k++) q[s4 + k] = turned[k] * l;
    const double nx = swept[1], ny = swept[2], nz = swept[3];
    double* v = w->velocity.data();
    const double into = v[s3] * nx + v[s3 + 1] * ny + v[s3 + 2] * nz;
    if (into < 0) {
      const double k = (1 + js_max(w->restitution[s], w->restitution[hit])) * into;
      v[s3] -= k * nx; v[s3 + 1] -= k * ny; v[s3 + 2] -= k * nz;
    }
    w->caught[s] = 1;
    hits[k] = 1;
  };
  if (threads) w->pool->run((int)fast_list.size(), one);
  else
    for (int k = 0; k < (int)fast_list.size(); k++) {
      one(k);
      if (rcs[k] != BM_OK) break;
    }
  w->device_lock = lock;
  for (int k = 0; k < (int)fast_list.size(); k++)
    if (rcs[k] != BM_OK) return rcs[k];
  int caught = 0;
  for (uint8_t h : hits) caught += h;
  if (w->broad.bodies_set)
    for (size_t k = 0; k < fast_list.size(); k++)
      if (hits[k]) w->broad.dirty.push_back(fast_list[k]);
  w->catches += caught;
  for (int i = 0; i < n; i++) w->fast[list[i]] = 0;
  if (caught == 0) return BM_OK;
  for (int m = 0; m < w->c.count; m++)
    if (w->caught[w->c.a[m]] == 1 || w->caught[w->c.b[m]] == 1) w->c.gap[m] = gap_of(w, m);
  for (int i = 0; i < n; i++) w->caught[list[i]] = 0;
  return BM_OK;
}

// src/contact.js:2178-2213 (touched): which pairs touched, and the begin or end of each that changed.
void touched(bm_world* w) {
  Contacts& c = w->c;
  Pairs& pr = w->pairs;
  // Each pair's rows are neighbours in narrow order: whether any of them meets (its gap within
  // tolerance, or a point pushed), and the first that does.
  auto first_hit = [&](int& f) {
    const int p = c.pair[c.row[f]];
    int hit = -1;
    for (; f < c.count && c.pair[c.row[f]] == p; f++) {
      const int m = c.row[f];
      if (hit >= 0) continue;
      bool meets = c.gap[m] <= c.tolerance[m];
      for (int k = 0; k < c.points[m] && !meets; k++) meets = c.max_normal_impulse[m * P + k] > 0;
      if (meets) hit = m;
    }
    return std::make_pair(p, hit);
  };
  auto event = [&](int p, int hit) {
    if (hit < 0) {
      w->events.end(pr.a[p], pr.b[p]);
      return;
    }
    // The deepest point, moved on by half of what the two bodies moved this step; the
    // fastest closing speed among the points.
    int deepest = hit * P;
    double speed = 0;
    for (int k = 0; k < c.points[hit]; k++) {
      const int i = hit * P + k;
      if (c.separation[i] < c.separation[deepest]) deepest = i;
      if (-c.relative_velocity[i] > speed) speed = -c.relative_velocity[i];
    }
    const int a3 = 3 * c.a[hit], b3 = 3 * c.b[hit], d3 = 3 * deepest, h3 = 3 * hit;
    const double* p3 = w->position.data();
    const double* st = w->start.data();
    double at[3];
    for (int k = 0; k < 3; k++)
      at[k] = c.point[d3 + k] + (c.shift[d3 + k] + p3[a3 + k] - st[a3 + k] + p3[b3 + k] - st[b3 + k]) * 0.5;
    w->events.begin(pr.a[p], pr.b[p], at[0], at[1], at[2], c.normal[h3], c.normal[h3 + 1], c.normal[h3 + 2], speed);
  };
  const int n = c.count;
  if (!w->pool || n < w->parallel_from) {
    for (int f = 0; f < n;) {
      const auto [p, hit] = first_hit(f);
      const uint8_t now = hit >= 0 ? 1 : 0;
      if (now == pr.touching[p]) continue;
      pr.touching[p] = now;
      event(p, hit);
    }
    return;
  }
  // A large world: the pairs' flags on the pool, in chunks cut between one pair's rows and the
  // next's (each pair's entries touched by one chunk alone); the pairs that changed listed per
  // chunk, and their events written after, in narrow order.
  const int chunk = 4096, chunks = (n + chunk - 1) / chunk;
  std::vector<std::vector<std::pair<int, int>>> changed((size_t)chunks);
  w->pool->run(chunks, [&](int k) {
    int f = k * chunk, e = std::min(n, f + chunk);
    if (k > 0)
      for (const int before = c.pair[c.row[f - 1]]; f < n && c.pair[c.row[f]] == before;) f++;
    if (e < n)
      for (const int last = c.pair[c.row[e - 1]]; e < n && c.pair[c.row[e]] == last;) e++;
    std::vector<std::pair<int, int>>& list = changed[k];
    list.clear();
    while (f < e) {
      const auto [p, hit] = first_hit(f);
      const uint8_t now = hit >= 0 ? 1 : 0;
      if (now == pr.touching[p]) continue;
      pr.touching[p] = now;
      list.emplace_back(p, hit);
    }
  });
  for (const auto& list : changed)
    for (const auto& [p, hit] : list) event(p, hit);
}

// src/contact.js:1067-1115 (store, keep) for level pairs: the device kept the other pairs'
// manifolds and worked out every manifold's slip; each solved level manifold goes into its
// pair's Memory, in the order narrow() found them, so the last one sets the count.
void store_memory(bm_world* w, int solved) {
  Contacts& c = w->c;
  Pairs& pr = w->pairs;
  auto one = [&](int f) {
    const int m = c.row[f];
    if (m >= solved || c.sub[m] < 0) return;
    const int sub = c.sub[m], m3 = 3 * m, s3 = 3 * sub;
    Memory& kept = pr.memory_of(c.pair[m]);
    kept.ensure(sub + 1);
    kept.count = sub + 1;
    for (int k = 0; k < 3; k++) kept.normal[s3 + k] = c.normal[m3 + k];
    kept.key[sub] = 0;
    const int count = c.points[m];
    kept.points[sub] = (uint8_t)count;
    const double t1x = c.t1[m3], t1y = c.t1[m3 + 1], t1z = c.t1[m3 + 2];
    const double t2x = c.t2[m3], t2y = c.t2[m3 + 1], t2z = c.t2[m3 + 2];
    const double r1 = c.roll_impulse1[m], r2 = c.roll_impulse2[m], sp = c.spin_impulse[m];
    kept.twist[s3] = t1x * r1 + t2x * r2 + c.normal[m3] * sp;
    kept.twist[s3 + 1] = t1y * r1 + t2y * r2 + c.normal[m3 + 1] * sp;
    kept.twist[s3 + 2] = t1z * r1 + t2z * r2 + c.normal[m3 + 2] * sp;
    const double l1 = c.tangent_impulse1[m], l2 = c.tangent_impulse2[m];
    kept.friction[s3] = t1x * l1 + t2x * l2;
    kept.friction[s3 + 1] = t1y * l1 + t2y * l2;
    kept.friction[s3 + 2] = t1z * l1 + t2z * l2;
    const double d1 = c.drift1[m], d2 = c.drift2[m];
    kept.slip[s3] = t1x * d1 + t2x * d2;
    kept.slip[s3 + 1] = t1y * d1 + t2y * d2;
    kept.slip[s3 + 2] = t1z * d1 + t2z * d2;
    for (int k = 0; k < count; k++) {
      kept.pid[P * sub + k] = c.pid[P * m + k];
      kept.normal_impulse[P * sub + k] = c.normal_impulse[P * m + k];
    }
  };
  const int n = c.count;
  if (!w->pool || n < w->parallel_from) {
    for (int f = 0; f < n; f++) one(f);
    return;
  }
  // A pair's rows are neighbours in narrow order (touched() reads them so too), and only they
  // write the pair's Memory: chunks cut where one pair's rows end and the next's begin write
  // disjoint Memories, each pair's rows still in narrow order.
  const int chunk = 4096, chunks = (n + chunk - 1) / chunk;
  w->pool->run(chunks, [&](int k) {
    int f = k * chunk, e = std::min(n, f + chunk);
    if (k > 0)
      for (const int before = c.pair[c.row[f - 1]]; f < n && c.pair[c.row[f]] == before;) f++;
    if (e < n)
      for (const int last = c.pair[c.row[e - 1]]; e < n && c.pair[c.row[e]] == last;) e++;
    for (; f < e; f++) one(f);
  });
}

// The device phase's columns for this world alone (src/world.js:956-1031 reads and writes
// them), into w->h.
void fill_columns(bm_world* w, double dt) {
  Contacts& c = w->c;
  Pairs& pr = w->pairs;
  c.ensure(c.count > 0 ? c.count : 1);
  pr.grow(pr.count > 0 ? pr.count : 1);
  w->h_body_offset[0] = 0; w->h_body_offset[1] = w->top;
  w->h_manifold_offset[0] = 0; w->h_manifold_offset[1] = c.count;
  w->h_pair_offset[0] = 0; w->h_pair_offset[1] = pr.count;
  w->h_awake_offset[0] = 0; w->h_awake_offset[1] = w->awake_count;
  for (int k = 0; k < BM_STARTS; k++) w->h_starts[k] = 0;
  w->h_solved = 0;
  w->h_substeps = w->substeps;
  w->h_dt = dt;
  bm_columns& h = w->h;
  h = bm_columns{};
  h.n_worlds = 1;
  h.body_offset = w->h_body_offset; h.manifold_offset = w->h_manifold_offset; h.pair_offset = w->h_pair_offset;
  h.awake_offset = w->h_awake_offset;
  h.starts = w->h_starts; h.solved = &w->h_solved; h.substeps = &w->h_substeps; h.dt = &w->h_dt; h.gravity = w->gravity;
  h.position = w->position.data(); h.rotation = w->rotation.data(); h.velocity = w->velocity.data();
  h.angular = w->angular.data(); h.start = w->start.data(); h.start_q = w->start_q.data(); h.inv_i = w->inv_i.data();
  h.tumbled = w->tumbled.data(); h.ceiling = w->ceiling.data(); h.dims = w->dims.data(); h.inv_mass = w->inv_mass.data();
  h.inertia = w->inertia.data(); h.inv_inertia = w->inv_inertia.data(); h.axes = w->axes.data();
  h.gravity_scale = w->gravity_scale.data(); h.damping = w->damping.data(); h.angular_damping = w->angular_damping.data();
  h.type = w->type.data(); h.shape = w->shape.data(); h.frozen = w->frozen.data(); h.lock = w->lock.data();
  h.framed = w->framed.data(); h.awake = w->awake.data(); h.joint_head = w->joint_head.data();
  h.awake_list = w->awake_list.data();
  h.a = c.a.data(); h.b = c.b.data(); h.pair = c.pair.data(); h.sub = c.sub.data(); h.row = c.row.data();
  h.colour = c.colour.data(); h.order = c.order.data(); h.points = c.points.data(); h.sliding = c.sliding.data();
  h.normal = c.normal.data(); h.t1 = c.t1.data(); h.t2 = c.t2.data(); h.friction = c.friction.data();
  h.restitution = c.restitution.data(); h.tolerance = c.tolerance.data(); h.patch = c.patch.data();
  h.roll_impulse1 = c.roll_impulse1.data(); h.roll_impulse2 = c.roll_impulse2.data(); h.spin_impulse = c.spin_impulse.data();
  h.tangent_impulse1 = c.tangent_impulse1.data(); h.tangent_impulse2 = c.tangent_impulse2.data();
  h.drift1 = c.drift1.data(); h.drift2 = c.drift2.data(); h.separation = c.separation.data(); h.point = c.point.data();
  h.shift = c.shift.data(); h.normal_impulse = c.normal_impulse.data(); h.max_normal_impulse = c.max_normal_impulse.data();
  h.relative_velocity = c.relative_velocity.data(); h.gap = c.gap.data(); h.now = c.now.data(); h.ra = c.ra.data();
  h.rb = c.rb.data(); h.la = c.la.data(); h.lb = c.lb.data(); h.adjusted = c.adjusted.data();
  h.normal_mass = c.normal_mass.data(); h.jac = c.jac.data(); h.block = c.block.data(); h.lever = c.lever.data();
  h.anchor_a = c.anchor_a.data(); h.anchor_b = c.anchor_b.data(); h.fjac = c.fjac.data();
  h.tangent_mass = c.tangent_mass.data(); h.roll_mass1 = c.roll_mass1.data(); h.roll_mass2 = c.roll_mass2.data();
  h.spin_mass = c.spin_mass.data(); h.pid = c.pid.data();
  h.pair_points = pr.points.data(); h.pair_pid = pr.pid.data(); h.pair_normal_impulse = pr.normal_impulse.data();
  h.pair_twist = pr.twist.data(); h.pair_friction = pr.friction.data(); h.pair_slip = pr.slip.data();
}

// ---- recording for the adjoint ----

uint64_t pair_key(int a, int b) {
  const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  return ((uint64_t)hi << 32) | lo;
}

// The state the held trajectory starts from, and the pairs' memory then as first rows.
void record_start(bm_world* w) {
  Record& R = w->record;
  R = Record{};
  R.on = true;
  R.top = w->top;
  R.catches = w->catches;
  R.row_offset.push_back(0);
  for (int s = 0; s < w->top; s++) {
    for (int k = 0; k < 3; k++) R.initial.push_back(w->position[3 * s + k]);
    for (int k = 0; k < 4; k++) R.initial.push_back(w->rotation[4 * s + k]);
    for (int k = 0; k < 3; k++) R.initial.push_back(w->velocity[3 * s + k]);
    for (int k = 0; k < 3; k++) R.initial.push_back(w->angular[3 * s + k]);
    for (int k = 0; k < 3; k++) R.left.push_back(w->tumbled[3 * s + k]);
    R.ceiling.push_back(w->ceiling[s]);
  }
  const Pairs& pr = w->pairs;
  for (int p = 0; p < pr.count; p++) {
    if (pr.points[p] == 0) continue;
    const int f = (int)R.first_count.size();
    R.first_count.push_back(pr.points[p]);
    for (int k = 0; k < P; k++) {
      R.first_pid.push_back(pr.pid[P * p + k]);
      R.first_impulse.push_back(pr.normal_impulse[P * p + k]);
    }
    for (int k = 0; k < 3; k++) {
      R.first_friction.push_back(pr.friction[3 * p + k]);
      R.first_twist.push_back(pr.twist[3 * p + k]);
      R.first_slip.push_back(pr.slip[3 * p + k]);
    }
    R.last[pair_key(pr.a[p], pr.b[p])] = -2 - f;
  }
}

// Before the phase: each found manifold held on its bodies (src/contact.js:540-570), and
// whether its pair carried a memory into it.
void record_before(bm_world* w, double dt) {
  Record& R = w->record;
  const Contacts& c = w->c;
  if (R.steps == 0) R.dt = dt;
  else if (dt != R.dt && R.broken.empty()) R.broken = "the recorded steps differ in dt";
  if (w->top != R.top && R.broken.empty()) R.broken = "bodies were added while recording";
  R.pre_normal.assign(3 * (size_t)c.count, 0);
  R.pre_ends.assign(6 * P * (size_t)c.count, 0);
  R.pre_shift.assign(3 * P * (size_t)c.count, 0);
  R.pre_had.assign(c.count, 0);
  const double* pos = w->position.data();
  const double* A = w->axes.data();
  for (int m = 0; m < c.count; m++) {
    if (c.sub[m] >= 0 && R.broken.empty()) R.broken = "a heightfield pair (the held step has no level pairs)";
    const int a = c.a[m], b = c.b[m], a3 = 3 * a, b3 = 3 * b, a9 = 9 * a, b9 = 9 * b;
    const double nx = c.normal[3 * m], ny = c.normal[3 * m + 1], nz = c.normal[3 * m + 2];
    double* ln = &R.pre_normal[3 * m];
    ln[0] = A[a9] * nx + A[a9 + 1] * ny + A[a9 + 2] * nz;
    ln[1] = A[a9 + 3] * nx + A[a9 + 4] * ny + A[a9 + 5] * nz;
    ln[2] = A[a9 + 6] * nx + A[a9 + 7] * ny + A[a9 + 8] * nz;
    for (int k = 0; k < c.points[m]; k++) {
      const int i = P * m + k;
      double* L = &R.pre_ends[6 * i];
      const double half = 0.5 * c.separation[i];
      const double* pt = &c.point[3 * i];
      const double* sh = &c.shift[3 * i];
      const double ax = pt[0] - nx * half - pos[a3], ay = pt[1] - ny * half - pos[a3 + 1], az = pt[2] - nz * half - pos[a3 + 2];
      const double bx = pt[0] + nx * half - pos[b3], by = pt[1] + ny * half - pos[b3 + 1],
                   bz = pt[2] + nz * half - pos[b3 + 2];
      for (int j = 0; j < 3; j++) R.pre_shift[3 * i + j] = sh[j];
      L[0] = A[a9] * ax + A[a9 + 1] * ay + A[a9 + 2] * az;
      L[1] = A[a9 + 3] * ax + A[a9 + 4] * ay + A[a9 + 5] * az;
      L[2] = A[a9 + 6] * ax + A[a9 + 7] * ay + A[a9 + 8] * az;
      L[3] = A[b9] * bx + A[b9 + 1] * by + A[b9 + 2] * bz;
      L[4] = A[b9 + 3] * bx + A[b9 + 4] * by + A[b9 + 5] * bz;
      L[5] = A[b9 + 6] * bx + A[b9 + 7] * by + A[b9 + 8] * bz;
      cored(w, a, L);
      cored(w, b, L + 3);
    }
    R.pre_had[m] = w->pairs.points[c.pair[m]] > 0 ? 1 : 0;
  }
}

// After the phase: the solved rows in row order, their colours and narrow order, and
// which row each pair's memory now comes from.
void record_after(bm_world* w) {
  Record& R = w->record;
  const Contacts& c = w->c;
  const int solved = w->h_solved, base = R.row_offset.back();
  for (int k = 0; k < BM_STARTS; k++) R.starts.push_back(w->h_starts[k]);
  for (int m = 0; m < c.count; m++)
    if (c.row[m] < solved) R.narrow.push_back(c.row[m]);
  for (int f = 0; f < solved; f++) {
    const int m = c.order[f];
    const int a = c.a[f], b = c.b[f];
    R.a.push_back(a); R.b.push_back(b);
    R.count.push_back(c.points[f]);
    for (int k = 0; k < P; k++) R.pid.push_back(c.pid[P * f + k]);
    for (int k = 0; k < 3; k++) R.local_normal.push_back(R.pre_normal[3 * m + k]);
    for (int k = 0; k < 6 * P; k++) R.ends.push_back(R.pre_ends[6 * P * m + k]);
    for (int k = 0; k < 3 * P; k++) R.shift.push_back(R.pre_shift[3 * P * m + k]);
    R.tolerance.push_back(c.tolerance[f]);
    R.patch.push_back(c.patch[f]);
    int src = -1;
    if (R.pre_had[m]) {
      auto it = R.last.find(pair_key(a, b));
      if (it == R.last.end()) {
        if (R.broken.empty()) R.broken = "a pair carried a memory the record does not hold";
      } else {
        src = it->second;
      }
    }
    R.src.push_back(src);
  }
  for (int f = 0; f < solved; f++) R.last[pair_key(c.a[f], c.b[f])] = base + f;
  R.row_offset.push_back(base + solved);
  R.steps++;
}

// The solved rows of the step just solved (laid out by colour, solved ones first) and the
// compulsory bytes of its soft step's sweeps, each field read or written once a pass and each
// body's fields once a manifold end (the gather and scatter every colour does). A manifold
// of n points, in bytes:
//   solve (bias, separations known): 394 fixed + 56 if n <= 2 + 8 n^2 if n >= 2 (the block)
//     + n (152 + 8 if n == 1) + bodies: v, w read 96 and written 96 (48 with one end static),
//     inverse masses 16, inverse inertias 96;
//   relax, and a step's first solve: 24 fewer fixed, 56 more a point, 160 more body bytes
//     (positions, starts, rotations);
//   warm start: 201 fixed + 16 if n <= 2 + 56 n + 304 body.
// A step of S substeps is S warm starts, S - 1 known solves, one first solve and S relaxes.
// `soft_ms` is the soft-step kernel's device time.
void solve_stats(bm_world* w, double soft_ms) {
  const Contacts& c = w->c;
  const double S = w->substeps;
  int64_t rows = 0;
  double bytes = 0;
  for (int f = 0; f < w->h_solved; f++) {
    const int n = c.points[f];
    rows += n;
    const bool both = w->inv_mass[c.a[f]] > 0 && w->inv_mass[c.b[f]] > 0;
    const double body = 96 + (both ? 96 : 48) + 16 + 96;
    const double fixed = 394 + (n <= 2 ? 56 : 0) + (n >= 2 ? 8.0 * n * n : 0);
    const double one = n == 1 ? 8 : 0;
    const double solve = fixed + n * (152 + one) + body;
    const double relax = (fixed - 24) + n * (152 + 56 + one) + body + 160;
    const double warm = 201 + (n <= 2 ? 16 : 0) + 56.0 * n + 304;
    bytes += S * warm + (S - 1) * solve + relax + S * relax;
  }
  w->step_rows = rows;
  w->step_bytes = bytes;
  w->step_soft_ms = soft_ms;
}

// The device phase: colour, rest, inertias .. store (src/world.js:956-1031), through the batch ABI.
int device_phase(bm_world* w, double dt) {
  fill_columns(w, dt);
  if (w->record.on) record_before(w, dt);
  bm_columns& h = w->h;
  if (w->tap) w->tap(w->tap_user, &h);
  bm_set_body_version(w->ctx, w->body_version);
  if (int rc = bm_upload(w->ctx, &h)) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
  if (int rc = bm_colour(w->ctx)) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
  if (int rc = bm_soft_step(w->ctx)) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
  if (int rc = bm_download(w->ctx, &h)) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
  if (w->broad.bodies_set) {
    // The step's state becomes the broadphase's on the device: every slot is as the host has it.
    if (int rc = bm_broad_keep(w->ctx)) return fail(w, rc, std::string("step: ") + bm_last_error(w->ctx));
    w->broad.state = true;
    w->broad.dirty.clear();
  }
  solve_stats(w, bm_last_kernel_ms(w->ctx));
  store_memory(w, w->h_solved);
  if (w->record.on) record_after(w);
  return BM_OK;
}

// ---- a step in phases (src/world.js:880-1072), so that worlds of a batch can share the
// device calls between them ----

// From the step's start to its first narrow pass's records (src/world.js:880-939).
int step_begin(bm_world* w, double dt) {
  w->error.clear();
  if (!(dt > 0) || dt == INFINITY) return fail(w, BM_ERR_ARGUMENT, "step: dt must be a positive number of seconds");
  ++w->step_count;
  w->events.count = 0;
  if (w->wake_pending) sync_pairs(w);
  // src/world.js:902-918: awake bodies' `previous` is where they start this step; a body that
  // fell asleep at the end of the last one moved in it, so it gets this once more, and its
  // stamp, once more.
  for (int i = 0; i < w->awake_count; i++) {
    const size_t t7 = 7 * (size_t)w->awake_list[i];
    for (int k = 0; k < 7; k++) w->previous[t7 + k] = w->transforms[t7 + k];
  }
  for (int i = 0; i < w->slept_count; i++) {
    const int s = w->slept[i];
    const size_t t7 = 7 * (size_t)s;
    for (int k = 0; k < 7; k++) w->previous[t7 + k] = w->transforms[t7 + k];
    w->moved_at[s] = w->step_count;
    log_moved(w, s);
  }
  w->slept_count = 0;
  w->step_dt = dt;
  const double gl = gravity_length(w);
  auto& bp = w->broad;
  bp.on_device = broad_on(w);
  bp.apart_ready = false;
  bool walked = false;
  if (bp.on_device) {
    // The world has begun its step: a device failure from here on leaves it part-way.
    if (int rc = broad_update(w, dt, gl)) return broke(w, rc);
    if (int rc = broad_candidates(w, walked)) return broke(w, rc);
  } else {
    for (int i = 0; i < w->awake_count; i++) bounds_of(w, w->awake_list[i], dt, gl);
    broadphase_update(w, dt);
  }
  if (!walked) candidates(w);
  w->c.count = 0;
  w->narrow_from = 0;
  if (w->wake_pending) sync_pairs(w);
  if (bp.on_device)
    if (int rc = broad_holds(w, 0)) return broke(w, rc);
  narrow_begin(w, 0);
  return BM_OK;
}

// After the device ran the batch: the pass goes on; when it ends and woke bodies, the next
// pass starts over the pairs they woke (src/world.js:940-946). True while records wait.
bool narrow_continue(bm_world* w, int& rc) {
  rc = BM_OK;
  if (narrow_resume(w, rc)) return true;
  if (rc != BM_OK) return false;
  w->narrow_from = w->pairs.active;
  if (!w->wake_pending) return false;
  sync_pairs(w);
  narrow_begin(w, w->narrow_from);
  return true;
}

// src/world.js:947-955: freezing, and the solver's step count.
void step_middle(bm_world* w) {
  if (w->sleeping && (w->frozen_count > 0 || w->still_ready)) freeze(w);
  if (w->c.count > 0) w->solver_steps++;
}

// src/world.js:1032-1072: the catch, touched, transforms, sleep.
int step_end(bm_world* w, double dt) {
  const int touching = w->c.count;
  w->strays_at = -1;
  if (int rc = catch_fast(w)) return rc;
  if (touching > 0) touched(w);
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (w->frozen[s] == 1) continue;
    if (w->frozen[s] == 2) w->frozen[s] = 1;
    write_transform(w, s);
    w->moved_at[s] = w->step_count;
  }
  w->step_dt = 0;
  if (w->sleeping) sleep_step(w, dt);
  return BM_OK;
}


}  // namespace

// Worlds stepped together: each world's order-defining host work on a thread of the pool,
// every narrowphase pass's records of all worlds in one device launch, and one colouring and
// soft-step launch for every world's contact phase.
struct bm_batch {
  bm_context* ctx = nullptr;
  std::string error;
  int broken = BM_OK;  // a step that failed part-way: every later step returns it
  std::string broken_msg;
  bool tables_stale = false;  // a world's hulls or fields changed since they were joined
  std::vector<bm_world*> worlds;
  std::mutex lock;
  Pool* pool = nullptr;
  // A merged meet pass: every world's body rows one after another (table), each world's pairs'
  // rows moved up by its first (ia, ib), and the pairs' numbers; then the results.
  std::vector<double, Unzeroed<double>> table, pair;
  std::vector<int32_t, Unzeroed<int32_t>> ia, ib;
  std::vector<double> normal, points, seps, shift, loss;
  std::vector<int32_t> count, flags, row;
  std::vector<uint32_t> ids;
  std::vector<int32_t> level_field, level_manifolds, level_count, level_start;
  std::vector<double> level_pos, level_axes, level_shape, level_box, level_margin, level_tol, level_normal,
      level_points, level_seps;
  std::vector<uint32_t> level_ids;
  int level_cap = 8;
  double times[BM_WORLD_TIMES] = {};  // as bm_world_times, for the batch's steps
};

namespace {

int batch_fail(bm_batch* B, int code, const std::string& m) {
  B->error = m;
  return code;
}

// Every world's waiting narrowphase records in one bm_meet_host and one bm_level_host call,
// the results handed back to each world as its own run_batch() would have left them.
int run_merged(bm_batch* B) {
  auto& W = B->worlds;
  const int nw = (int)W.size();
  std::vector<int> at(nw + 1, 0);
  for (int i = 0; i < nw; i++) at[i + 1] = at[i] + (int)W[i]->batch_pair.size() / 6;
  const int n = at[nw];
  if (n > 0) {
    // Each world's bodies that meet, once each; then their rows, the pairs' rows moved up by
    // the world's first, and the pairs' numbers, each world into its place.
    std::vector<int> bt(nw + 1, 0);
    B->pool->run(nw, [&](int i) {
      bm_world* w = W[i];
      bt[i + 1] = at[i + 1] > at[i] ? meet_rows(w, at[i + 1] - at[i]) : 0;
    });
    for (int i = 0; i < nw; i++) bt[i + 1] += bt[i];
    B->table.resize(32 * (size_t)bt[nw]); B->ia.resize(n); B->ib.resize(n); B->pair.resize(6 * (size_t)n);
    B->pool->run(nw, [&](int i) {
      bm_world* w = W[i];
      const int k = at[i + 1] - at[i];
      if (k == 0) return;
      const std::vector<int32_t>& list = w->meet_bodies;
      for (size_t r = 0; r < list.size(); r++) load(w, list[r], &B->table[32 * ((size_t)bt[i] + r)]);
      for (int j = 0; j < k; j++) {
        B->ia[at[i] + j] = bt[i] + w->meet_ta[j];
        B->ib[at[i] + j] = bt[i] + w->meet_tb[j];
      }
      std::copy(w->batch_pair.begin(), w->batch_pair.begin() + 6 * (size_t)k, B->pair.begin() + 6 * (size_t)at[i]);
    });
    B->count.resize(n); B->flags.resize(n); B->row.resize(n); B->normal.resize(3 * (size_t)n);
    B->points.resize(12 * (size_t)n); B->seps.resize(4 * (size_t)n); B->ids.resize(4 * (size_t)n);
    B->shift.resize(12 * (size_t)n); B->loss.resize(n);
    if (int rc = bm_meet_packed(B->ctx, bt[nw], B->table.data(), n, B->ia.data(), B->ib.data(), B->pair.data(),
                                B->count.data(), B->flags.data(), B->row.data(), B->normal.data(), B->points.data(),
                                B->seps.data(), B->ids.data(), B->shift.data(), B->loss.data()))
      return batch_fail(B, rc, std::string("step: ") + bm_last_error(B->ctx));
    B->pool->run(nw, [&](int i) {
      bm_world* w = W[i];
      const int k0 = at[i], k = at[i + 1] - at[i];
      w->batch_ia.clear(); w->batch_ib.clear(); w->batch_pair.clear();
      if (k == 0) return;
      // After the results of the world's pass's earlier runs: its pairs' counts and flags, and
      // their packed rows (the batch's rows are in pair order, so a world's are one run).
      const size_t b = w->meet_base, e = b + k, rb = w->meet_found;
      w->batch_count.resize(e);
      w->batch_flags.resize(e);
      w->batch_row.resize(e);
      std::copy(B->count.begin() + k0, B->count.begin() + k0 + k, w->batch_count.begin() + b);
      std::copy(B->flags.begin() + k0, B->flags.begin() + k0 + k, w->batch_flags.begin() + b);
      int32_t lo = -1, hi = -1;
      for (int j = 0; j < k; j++) {
        const int32_t r = B->row[k0 + j];
        if (r >= 0) {
          if (lo < 0) lo = r;
          hi = r + 1;
        }
        w->batch_row[b + j] = r >= 0 ? r - lo + (int32_t)rb : -1;
      }
      const size_t rows = lo >= 0 ? (size_t)(hi - lo) : 0;
      auto put = [&](auto& to, const auto& from, size_t width) {
        to.resize(width * (rb + rows));
        if (rows > 0) std::copy(from.begin() + width * lo, from.begin() + width * hi, to.begin() + width * rb);
      };
      put(w->batch_normal, B->normal, 3);
      put(w->batch_points, B->points, 12);
      put(w->batch_seps, B->seps, 4);
      put(w->batch_ids, B->ids, 4);
      put(w->batch_shift, B->shift, 12);
      put(w->batch_loss, B->loss, 1);
      w->meet_base = e;
      w->meet_found = rb + rows;
    });
  }
  std::vector<int> lat(nw + 1, 0);
  for (int i = 0; i < nw; i++) lat[i + 1] = lat[i] + (int)W[i]->level_field.size();
  const int ln = lat[nw];
  if (ln > 0) {
    for (auto* v : {&B->level_pos, &B->level_axes, &B->level_shape, &B->level_box, &B->level_margin, &B->level_tol})
      v->clear();
    B->level_field.clear();
    for (bm_world* w : W) {
      B->level_field.insert(B->level_field.end(), w->level_field.begin(), w->level_field.end());
      B->level_pos.insert(B->level_pos.end(), w->level_pos.begin(), w->level_pos.end());
      B->level_axes.insert(B->level_axes.end(), w->level_axes.begin(), w->level_axes.end());
      B->level_shape.insert(B->level_shape.end(), w->level_shape.begin(), w->level_shape.end());
      B->level_box.insert(B->level_box.end(), w->level_box.begin(), w->level_box.end());
      B->level_margin.insert(B->level_margin.end(), w->level_margin.begin(), w->level_margin.end());
      B->level_tol.insert(B->level_tol.end(), w->level_tol.begin(), w->level_tol.end());
    }
    for (;;) {
      const size_t slots = (size_t)ln * B->level_cap;
      B->level_manifolds.resize(ln); B->level_start.resize(ln + 1); B->level_count.resize(slots);
      B->level_normal.resize(3 * slots); B->level_points.resize(12 * slots); B->level_seps.resize(4 * slots);
      B->level_ids.resize(4 * slots);
      if (int rc = bm_level_packed(B->ctx, ln, B->level_field.data(), B->level_pos.data(), B->level_axes.data(),
                                   B->level_shape.data(), B->level_box.data(), B->level_margin.data(),
                                   B->level_tol.data(), B->level_cap, B->level_manifolds.data(),
                                   B->level_start.data(), B->level_count.data(), B->level_normal.data(),
                                   B->level_points.data(), B->level_seps.data(), B->level_ids.data()))
        return batch_fail(B, rc, std::string("step: ") + bm_last_error(B->ctx));
      int most = 0;
      for (int k = 0; k < ln; k++) most = std::max(most, (int)B->level_manifolds[k]);
      if (most <= B->level_cap) break;
      B->level_cap = most;
    }
    const size_t cap = B->level_cap;
    for (int i = 0; i < nw; i++) {
      bm_world* w = W[i];
      const size_t k0 = lat[i], k = lat[i + 1] - lat[i];
      for (auto* v : {&w->level_pos, &w->level_axes, &w->level_shape, &w->level_box, &w->level_margin, &w->level_tol})
        v->clear();
      w->level_field.clear();
      if (k == 0) continue;
      w->level_cap = (int)cap;
      // The world's pairs and their packed rows, after its pass's earlier runs', the starts
      // counted from the world's first row.
      const size_t b = w->level_base, wr = b > 0 ? (size_t)w->level_start[b] : 0;
      const size_t r0 = B->level_start[k0], r1 = B->level_start[k0 + k], rows = r1 - r0;
      w->level_manifolds.resize(b + k);
      std::copy(B->level_manifolds.begin() + k0, B->level_manifolds.begin() + k0 + k, w->level_manifolds.begin() + b);
      w->level_start.resize(b + k + 1);
      for (size_t j = 0; j <= k; j++) w->level_start[b + j] = B->level_start[k0 + j] - (int32_t)r0 + (int32_t)wr;
      auto put = [&](auto& to, const auto& from, size_t width) {
        to.resize(width * (wr + rows));
        std::copy(from.begin() + width * r0, from.begin() + width * r1, to.begin() + width * wr);
      };
      put(w->level_count, B->level_count, 1);
      put(w->level_normal, B->level_normal, 3);
      put(w->level_points, B->level_points, 12);
      put(w->level_seps, B->level_seps, 4);
      put(w->level_ids, B->level_ids, 4);
      w->level_base = b + k;
    }
  }
  return BM_OK;
}

// Every world's hulls and fields joined into the device's tables, each world's numbers moved
// up by the hulls and fields before it.
int join_tables(bm_batch* B) {
  std::vector<int32_t> hv{0}, hf{0}, hl{0}, st, lp, cols, rows;
  std::vector<double> vt, pl, sp, heights;
  std::vector<int64_t> off;
  for (bm_world* w : B->worlds) {
    w->hull_base = (int)hv.size() - 1;
    const int nh = w->given_hull_v.empty() ? 0 : (int)w->given_hull_v.size() - 1;
    for (int h = 0; h < nh; h++) {
      hv.push_back(hv.back() + w->given_hull_v[h + 1] - w->given_hull_v[h]);
      hf.push_back(hf.back() + w->given_hull_f[h + 1] - w->given_hull_f[h]);
      hl.push_back(hl.back() + w->given_hull_l[h + 1] - w->given_hull_l[h]);
    }
    vt.insert(vt.end(), w->given_vertices.begin(), w->given_vertices.end());
    pl.insert(pl.end(), w->given_planes.begin(), w->given_planes.end());
    st.insert(st.end(), w->given_start.begin(), w->given_start.end());
    lp.insert(lp.end(), w->given_loops.begin(), w->given_loops.end());
    w->field_base = (int)cols.size();
    for (size_t f = 0; f < w->given_columns.size(); f++) {
      cols.push_back(w->given_columns[f]);
      rows.push_back(w->given_rows[f]);
      sp.push_back(w->given_spacing[f]);
      off.push_back((int64_t)heights.size() + w->given_offsets[f]);
    }
    heights.insert(heights.end(), w->given_heights.begin(), w->given_heights.end());
  }
  const int nh = (int)hv.size() - 1;
  if (int rc = bm_set_hulls(B->ctx, nh, hv.data(), hf.data(), hl.data(), vt.data(), st.data(), lp.data(), pl.data()))
    return batch_fail(B, rc, std::string("step: ") + bm_last_error(B->ctx));
  if (int rc = bm_set_fields(B->ctx, (int)cols.size(), cols.data(), rows.data(), sp.data(), off.data(), heights.data()))
    return batch_fail(B, rc, std::string("step: ") + bm_last_error(B->ctx));
  B->tables_stale = false;
  return BM_OK;
}

// The first world whose phase failed, as the batch's error.
int first_failure(bm_batch* B, const std::vector<int>& rcs) {
  for (size_t i = 0; i < rcs.size(); i++)
    if (rcs[i] != BM_OK)
      return batch_fail(B, rcs[i], "world " + std::to_string(i) + ": " + B->worlds[i]->error);
  return BM_OK;
}

}  // namespace

extern "C" {

// A number no world's bodies have had: names a world's constant body columns after each add.
uint64_t next_body_version() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1);
}

int bm_world_create(bm_world** out, int32_t device, const double* gravity, int32_t substeps, int32_t sleeping) {
  if (!out) return BM_ERR_ARGUMENT;
  *out = nullptr;
  bm_world* w = new bm_world();
  if (int rc = bm_create(&w->ctx, device)) {
    delete w;
    return rc;
  }
  // The host never reads the soft step's own columns between steps.
  bm_set_device_only(w->ctx, 1);
  if (gravity)
    for (int k = 0; k < 3; k++) w->gravity[k] = gravity[k];
  w->substeps = substeps > 0 ? substeps : 8;
  w->sleeping = sleeping != 0;
  *out = w;
  return BM_OK;
}

int bm_world_destroy(bm_world* w) {
  if (!w) return BM_OK;
  if (!w->owns_ctx) return fail(w, BM_ERR_ARGUMENT, "bm_world_destroy: the world belongs to a batch (bm_batch_destroy)");
  delete w->pool;
  bm_destroy(w->ctx);
  delete w;
  return BM_OK;
}

const char* bm_world_error(const bm_world* w) { return w ? w->error.c_str() : "bm_world: no world"; }

int bm_world_add(bm_world* w, const bm_bodies* b) {
  if (!w || !b) return BM_ERR_ARGUMENT;
  w->error.clear();
  w->broad.bodies_set = false;  // the device's bodies are given again at the next step
  w->body_version = next_body_version();
  if (b->n < 0) return fail(w, BM_ERR_ARGUMENT, "addBody: n must be 0 or more");
  for (int i = 0; i < b->n; i++) {
    if (!b->type || !b->shape || !b->lock || !b->position || !b->rotation || !b->velocity || !b->angular || !b->dims ||
        !b->inertia || !b->inv_inertia || !b->mass || !b->inv_mass || !b->gravity_scale || !b->damping ||
        !b->angular_damping || !b->size || !b->reach || !b->friction || !b->restitution)
      return fail(w, BM_ERR_ARGUMENT, "addBody: every column is needed");
    if (b->shape[i] != BOX && b->shape[i] != SPHERE && b->shape[i] != CAPSULE && b->shape[i] != HULL &&
        b->shape[i] != HEIGHTFIELD)
      return fail(w, BM_ERR_UNSUPPORTED, "addBody: this world takes boxes, spheres, capsules, hulls and heightfields");
    if (b->shape[i] == HEIGHTFIELD && (!b->field || b->field[i] < 0 || b->field[i] >= (int)w->fields.size()))
      return fail(w, BM_ERR_ARGUMENT, "addBody: a heightfield body names no field of the table (bm_world_set_fields)");
    if (b->shape[i] == HEIGHTFIELD && b->type[i] != STATIC)
      return fail(w, BM_ERR_ARGUMENT, "addBody: level geometry is never dynamic");
    if (b->shape[i] == HULL && (!b->hull || b->hull[i] < 0 || b->hull[i] >= (int)w->hulls.size()))
      return fail(w, BM_ERR_ARGUMENT, "addBody: a hull body names no hull of the table (bm_world_set_hulls)");
    const bool framed = b->framed && b->framed[i] != 0;
    if (framed && (!b->offset || !b->frame))
      return fail(w, BM_ERR_ARGUMENT, "addBody: a framed body needs its offset and frame");
    if (framed && b->shape[i] != HULL)
      return fail(w, BM_ERR_UNSUPPORTED, "addBody: a framed box, sphere or capsule has parts, which this world does not take");
    if (b->type[i] != DYNAMIC && b->type[i] != STATIC)
      return fail(w, BM_ERR_UNSUPPORTED, "addBody: this world takes dynamic and static bodies");
  }
  const int first = w->top, top = w->top + b->n;
  auto grow = [&](auto& v, int per) { v.resize((size_t)per * top); };
  for (auto* v : {&w->position, &w->velocity, &w->angular, &w->dims, &w->inertia, &w->inv_inertia, &w->start, &w->tumbled})
    grow(*v, 3);
  for (auto* v : {&w->rotation, &w->start_q, &w->frame}) grow(*v, 4);
  grow(w->offset, 3); grow(w->hull, 1); grow(w->field, 1);
  grow(w->inv_i, 6); grow(w->axes, 9); grow(w->box, 6); grow(w->rest, 7);
  for (auto* v : {&w->mass, &w->inv_mass, &w->motion, &w->ceiling, &w->gravity_scale, &w->damping, &w->angular_damping,
                  &w->size, &w->reach, &w->friction, &w->restitution, &w->still, &w->least})
    grow(*v, 1);
  for (auto* v : {&w->type, &w->shape, &w->frozen, &w->calm, &w->lock, &w->framed, &w->awake, &w->fast, &w->caught,
                  &w->is_moved, &w->fresh})
    grow(*v, 1);
  for (auto* v : {&w->joint_head, &w->in_awake, &w->awake_list, &w->island, &w->island_next, &w->parent, &w->slept,
                  &w->leaf})
    grow(*v, 1);
  w->which.resize(top, -1);
  w->transforms.resize(7 * (size_t)top);
  w->previous.resize(7 * (size_t)top);
  w->moved_at.resize(top);
  w->seen.resize(top);
  w->top = top;
  for (int i = 0; i < b->n; i++) {
    const int s = first + i;
    for (int k = 0; k < 3; k++) {
      w->position[3 * s + k] = b->position[3 * i + k];
      w->velocity[3 * s + k] = b->velocity[3 * i + k];
      w->angular[3 * s + k] = b->angular[3 * i + k];
      w->dims[3 * s + k] = b->dims[3 * i + k];
      w->inertia[3 * s + k] = b->inertia[3 * i + k];
      w->inv_inertia[3 * s + k] = b->inv_inertia[3 * i + k];
      w->tumbled[3 * s + k] = NAN;  // never equal to any spin (src/world.js:2260)
    }
    for (int k = 0; k < 4; k++) w->rotation[4 * s + k] = b->rotation[4 * i + k];
    w->mass[s] = b->mass[i]; w->inv_mass[s] = b->inv_mass[i];
    w->gravity_scale[s] = b->gravity_scale[i]; w->damping[s] = b->damping[i];
    w->angular_damping[s] = b->angular_damping[i]; w->size[s] = b->size[i]; w->reach[s] = b->reach[i];
    w->friction[s] = b->friction[i]; w->restitution[s] = b->restitution[i];
    w->type[s] = b->type[i]; w->shape[s] = b->shape[i]; w->lock[s] = b->lock[i];
    w->hull[s] = b->shape[i] == HULL ? b->hull[i] : -1;
    w->field[s] = b->shape[i] == HEIGHTFIELD ? b->field[i] : -1;
    w->framed[s] = b->framed && b->framed[i] != 0 ? 1 : 0;
    for (int k = 0; k < 3; k++) w->offset[3 * s + k] = w->framed[s] ? b->offset[3 * i + k] : 0;
    for (int k = 0; k < 4; k++) w->frame[4 * s + k] = w->framed[s] ? b->frame[4 * i + k] : (k == 3 ? 1 : 0);
    // addBody (src/world.js:563-578): placed (transforms and previous, src/world.js:2283-2295),
    // bounded, filed, awake unless static, stamped as moved since the last step (1450-1454).
    write_transform(w, s);
    for (int k = 0; k < 7; k++) w->previous[7 * (size_t)s + k] = w->transforms[7 * (size_t)s + k];
    for (int k = 0; k < 3; k++) w->start[3 * s + k] = w->position[3 * s + k];
    bounds_of(w, s, 0, 0);
    broadphase_add(w, s);
    if (w->type[s] != STATIC) awaken(w, s);
    w->moved_at[s] = w->step_count + 1;
    log_moved(w, s);
  }
  return BM_OK;
}

int bm_world_set_hulls(bm_world* w, int32_t n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l,
                       const double* vertices, const int32_t* start, const int32_t* loops, const double* planes) {
  if (w) w->broad.bodies_set = false;
  if (!w) return BM_ERR_ARGUMENT;
  w->error.clear();
  for (int s = 0; s < w->top; s++)
    if (w->shape[s] == HULL && w->hull[s] >= n)
      return fail(w, BM_ERR_ARGUMENT, "bm_world_set_hulls: a body already added names a hull the table lacks");
  // The device's copy first: it checks the table's layout and limits. A batch's worlds keep
  // theirs, and the batch joins them before its next step.
  {
    std::unique_lock<std::mutex> held;
    if (w->device_lock) held = std::unique_lock<std::mutex>(*w->device_lock);
    if (int rc = bm_set_hulls(w->ctx, n, hull_v, hull_f, hull_l, vertices, start, loops, planes))
      return fail(w, rc, bm_last_error(w->ctx));
  }
  if (w->tables_stale) {
    *w->tables_stale = true;
    w->given_hull_v.assign(hull_v, hull_v + n + 1);
    w->given_hull_f.assign(hull_f, hull_f + n + 1);
    w->given_hull_l.assign(hull_l, hull_l + n + 1);
    w->given_vertices.assign(vertices, vertices + 3 * (size_t)(n ? hull_v[n] : 0));
    w->given_planes.assign(planes, planes + 4 * (size_t)(n ? hull_f[n] : 0));
    w->given_start.assign(start, start + (size_t)(n ? hull_f[n] : 0));
    w->given_loops.assign(loops, loops + (size_t)(n ? hull_l[n] : 0));
  }
  std::vector<bm_world::HostHull> hulls(n);
  std::vector<double> hv;
  for (int h = 0; h < n; h++) {
    auto& H = hulls[h];
    H.v0 = (int)(hv.size() / 3);
    H.V = hull_v[h + 1] - hull_v[h];
    const double* v = vertices + 3 * (size_t)hull_v[h];
    hv.insert(hv.end(), v, v + 3 * H.V);
    // src/hull.js:63-83: its box in its own frame, and its thinnest extent across a face.
    for (int k = 0; k < 3; k++) { H.lo[k] = INFINITY; H.hi[k] = -INFINITY; }
    for (int i = 0; i < H.V; i++)
      for (int k = 0; k < 3; k++) { H.lo[k] = js_min(H.lo[k], v[3 * i + k]); H.hi[k] = js_max(H.hi[k], v[3 * i + k]); }
    double size = INFINITY;
    const double* p = planes + 4 * (size_t)hull_f[h];
    for (int f = 0; f < hull_f[h + 1] - hull_f[h]; f++) {
      const double nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
      double low = INFINITY;
      for (int i = 0; i < H.V; i++) low = js_min(low, nx * v[3 * i] + ny * v[3 * i + 1] + nz * v[3 * i + 2]);
      size = js_min(size, p[4 * f + 3] - low);
    }
    H.size = size;
  }
  w->hulls.swap(hulls);
  w->hull_vertices.swap(hv);
  return BM_OK;
}

int bm_world_set_fields(bm_world* w, int32_t n, const int32_t* columns, const int32_t* rows, const double* spacing,
                        const int64_t* offsets, const double* heights) {
  if (w) w->broad.bodies_set = false;
  if (!w) return BM_ERR_ARGUMENT;
  w->error.clear();
  for (int s = 0; s < w->top; s++)
    if (w->shape[s] == HEIGHTFIELD && w->field[s] >= n)
      return fail(w, BM_ERR_ARGUMENT, "bm_world_set_fields: a body already added names a field the table lacks");
  // The device's copy first: it checks the table's layout. A batch joins its worlds' tables.
  {
    std::unique_lock<std::mutex> held;
    if (w->device_lock) held = std::unique_lock<std::mutex>(*w->device_lock);
    if (int rc = bm_set_fields(w->ctx, n, columns, rows, spacing, offsets, heights))
      return fail(w, rc, bm_last_error(w->ctx));
  }
  if (w->tables_stale) {
    *w->tables_stale = true;
    w->given_columns.assign(columns, columns + n);
    w->given_rows.assign(rows, rows + n);
    w->given_spacing.assign(spacing, spacing + n);
    w->given_offsets.assign(offsets, offsets + n);
    w->given_heights.clear();
    for (int f = 0; f < n; f++) {
      const double* h = heights + offsets[f];
      w->given_offsets[f] = (int64_t)w->given_heights.size();
      w->given_heights.insert(w->given_heights.end(), h, h + (size_t)columns[f] * rows[f]);
    }
  }
  std::vector<bm_world::HostField> fields(n);
  std::vector<double> all;
  for (int f = 0; f < n; f++) {
    auto& F = fields[f];
    F.columns = columns[f]; F.rows = rows[f]; F.spacing = spacing[f];
    F.offset = all.size();
    const double* h = heights + offsets[f];
    const size_t count = (size_t)columns[f] * rows[f];
    all.insert(all.end(), h, h + count);
    // src/mesh.js:296-305: centred on its body; its box and reach.
    F.x0 = -((F.columns - 1) * F.spacing) / 2;
    F.z0 = -((F.rows - 1) * F.spacing) / 2;
    double lo = INFINITY, hi = -INFINITY;
    for (size_t k = 0; k < count; k++) { lo = js_min(lo, h[k]); hi = js_max(hi, h[k]); }
    F.lo[0] = F.x0; F.lo[1] = lo; F.lo[2] = F.z0;
    F.hi[0] = -F.x0; F.hi[1] = hi; F.hi[2] = -F.z0;
    F.reach = std::sqrt(F.x0 * F.x0 + F.z0 * F.z0 + js_max(lo * lo, hi * hi));
  }
  w->fields.swap(fields);
  w->field_heights.swap(all);
  return BM_OK;
}

// src/world.js:880-1072 (step), for the bodies above.
int bm_world_step(bm_world* w, double dt) {
  if (!w) return BM_ERR_ARGUMENT;
  if (!w->owns_ctx) return fail(w, BM_ERR_ARGUMENT, "step: the world steps with its batch (bm_batch_step)");
  using clock = std::chrono::steady_clock;
  auto t = clock::now();
  auto lap = [&](int k) {
    const auto now = clock::now();
    w->times[k] += std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
  };
  if (w->broken != BM_OK) return fail(w, w->broken, w->broken_msg);
  // A step that fails before anything changed (dt) leaves the world as it was; one that
  // fails later leaves it half stepped, and every later step is refused.
  if (int rc = step_begin(w, dt)) return rc;
  lap(0);
  int rc = BM_OK;
  for (;;) {
    if ((rc = run_batch(w)) != BM_OK) return broke(w, rc);
    lap(1);
    const bool more = narrow_continue(w, rc);
    lap(2);
    if (!more) break;
  }
  if (rc != BM_OK) return broke(w, rc);
  step_middle(w);
  lap(3);
  if ((rc = device_phase(w, dt)) != BM_OK) return broke(w, rc);
  lap(4);
  rc = step_end(w, dt);
  lap(5);
  return rc != BM_OK ? broke(w, rc) : rc;
}

int bm_world_times(const bm_world* w, double* out) {
  if (!w || !out) return BM_ERR_ARGUMENT;
  for (int k = 0; k < BM_WORLD_TIMES; k++) out[k] = w->times[k];
  return BM_OK;
}

int bm_world_tap(bm_world* w, bm_columns_tap tap, void* user) {
  if (!w) return BM_ERR_ARGUMENT;
  w->tap = tap;
  w->tap_user = user;
  return BM_OK;
}

int bm_batch_create(bm_batch** out, int32_t device, int32_t threads) {
  if (!out) return BM_ERR_ARGUMENT;
  *out = nullptr;
  if (threads < 0) return BM_ERR_ARGUMENT;
  bm_batch* B = new bm_batch();
  if (int rc = bm_create(&B->ctx, device)) {
    delete B;
    return rc;
  }
  bm_set_device_only(B->ctx, 1);
  bm_set_graph(B->ctx, 1);
  B->pool = new Pool(threads > 1 ? threads - 1 : 0);
  *out = B;
  return BM_OK;
}

int bm_batch_destroy(bm_batch* B) {
  if (!B) return BM_OK;
  delete B->pool;
  for (bm_world* w : B->worlds) delete w;
  bm_destroy(B->ctx);
  delete B;
  return BM_OK;
}

const char* bm_batch_error(const bm_batch* B) { return B ? B->error.c_str() : "bm_batch: no batch"; }

int bm_batch_world(bm_batch* B, bm_world** out, const double* gravity, int32_t substeps, int32_t sleeping) {
  if (!B || !out) return BM_ERR_ARGUMENT;
  B->error.clear();
  bm_world* w = new bm_world();
  w->ctx = B->ctx;
  w->owns_ctx = false;
  w->device_lock = &B->lock;
  w->tables_stale = &B->tables_stale;
  if (gravity)
    for (int k = 0; k < 3; k++) w->gravity[k] = gravity[k];
  w->substeps = substeps > 0 ? substeps : 8;
  w->sleeping = sleeping != 0;
  B->worlds.push_back(w);
  *out = w;
  return BM_OK;
}

// Every world of the batch is left half stepped by a failure after the worlds began their
// step: the batch, and each of its worlds, keep it.
int batch_broke(bm_batch* B, int code) {
  B->broken = code;
  B->broken_msg = B->error + " (the batch's step stopped part-way; its worlds take no more steps)";
  for (bm_world* w : B->worlds) {
    if (w->broken != BM_OK) continue;
    w->error = B->broken_msg;
    broke(w, code);
  }
  return code;
}

int bm_batch_step(bm_batch* B, double dt) {
  if (!B) return BM_ERR_ARGUMENT;
  B->error.clear();
  if (B->broken != BM_OK) return batch_fail(B, B->broken, B->broken_msg);
  auto& W = B->worlds;
  const int n = (int)W.size();
  if (n == 0) return BM_OK;
  using clock = std::chrono::steady_clock;
  auto t = clock::now();
  auto lap = [&](int k) {
    const auto now = clock::now();
    B->times[k] += std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
  };
  if (B->tables_stale) {
    if (int rc = join_tables(B)) return rc;
  }
  if (!(dt > 0) || dt == INFINITY) return batch_fail(B, BM_ERR_ARGUMENT, "step: dt must be a positive number of seconds");
  std::vector<int> rcs(n, BM_OK);
  B->pool->run(n, [&](int i) { rcs[i] = step_begin(W[i], dt); });
  if (int rc = first_failure(B, rcs)) return batch_broke(B, rc);
  lap(0);
  std::vector<uint8_t> active(n, 1);
  for (;;) {
    if (int rc = run_merged(B)) return batch_broke(B, rc);
    lap(1);
    B->pool->run(n, [&](int i) {
      if (!active[i]) return;
      int rc = BM_OK;
      active[i] = narrow_continue(W[i], rc) ? 1 : 0;
      rcs[i] = rc;
    });
    if (int rc = first_failure(B, rcs)) return batch_broke(B, rc);
    bool any = false;
    for (int i = 0; i < n; i++) any = any || active[i];
    lap(2);
    if (!any) break;
  }
  std::vector<bm_columns*> hs(n);
  B->pool->run(n, [&](int i) {
    step_middle(W[i]);
    fill_columns(W[i], dt);
    if (W[i]->record.on) record_before(W[i], dt);
    hs[i] = &W[i]->h;
  });
  lap(3);
  std::vector<const bm_columns*> ins(hs.begin(), hs.end());
  auto device_failed = [&](int rc) {
    batch_fail(B, rc, std::string("step: ") + bm_last_error(B->ctx));
    return batch_broke(B, rc);
  };
  // The batch's bodies' constant columns are named by its worlds' versions, in order.
  uint64_t version = 1469598103934665603ull;
  for (int i = 0; i < n; i++) version = (version ^ W[i]->body_version) * 1099511628211ull;
  bm_set_body_version(B->ctx, version != 0 ? version : 1);
  if (int rc = bm_upload_worlds(B->ctx, n, ins.data())) return device_failed(rc);
  if (int rc = bm_contact_phase(B->ctx)) return device_failed(rc);
  if (int rc = bm_download_worlds(B->ctx, n, hs.data())) return device_failed(rc);
  lap(4);
  const double soft_ms = bm_last_kernel_ms(B->ctx);
  B->pool->run(n, [&](int i) {
    solve_stats(W[i], soft_ms);
    store_memory(W[i], W[i]->h_solved);
    if (W[i]->record.on) record_after(W[i]);
    rcs[i] = step_end(W[i], dt);
  });
  lap(5);
  if (int rc = first_failure(B, rcs)) return batch_broke(B, rc);
  return BM_OK;
}

int bm_batch_set_graph(bm_batch* B, int32_t on) {
  if (!B) return BM_ERR_ARGUMENT;
  return bm_set_graph(B->ctx, on);
}

int bm_batch_set_launch(bm_batch* B, int32_t shape, int32_t resident) {
  if (!B) return BM_ERR_ARGUMENT;
  const int rc = bm_set_launch(B->ctx, shape, resident);
  if (rc != BM_OK) batch_fail(B, rc, bm_last_error(B->ctx));
  return rc;
}

int bm_batch_graph_counts(const bm_batch* B, int64_t* out) {
  if (!B || !out) return BM_ERR_ARGUMENT;
  return bm_graph_counts(B->ctx, out);
}

int bm_batch_times(const bm_batch* B, double* out) {
  if (!B || !out) return BM_ERR_ARGUMENT;
  for (int k = 0; k < BM_WORLD_TIMES; k++) out[k] = B->times[k];
  return BM_OK;
}

int bm_world_read(const bm_world* w, bm_world_state* out) {
  if (!w || !out) return BM_ERR_ARGUMENT;
  out->top = w->top;
  out->awake = w->awake_count;
  out->contacts = w->c.count;
  out->pairs = w->pairs.count;
  out->frozen = w->frozen_count;
  out->impacts = w->impacts;
  out->catches = w->catches;
  out->rows = w->step_rows;
  out->soft_bytes = w->step_bytes;
  out->soft_ms = w->step_soft_ms;
  if (out->transforms) std::memcpy(out->transforms, w->transforms.data(), 7 * sizeof(float) * w->top);
  if (out->previous) std::memcpy(out->previous, w->previous.data(), 7 * sizeof(float) * w->top);
  if (out->position) std::memcpy(out->position, w->position.data(), 3 * sizeof(double) * w->top);
  if (out->rotation) std::memcpy(out->rotation, w->rotation.data(), 4 * sizeof(double) * w->top);
  if (out->velocity) std::memcpy(out->velocity, w->velocity.data(), 3 * sizeof(double) * w->top);
  if (out->angular) std::memcpy(out->angular, w->angular.data(), 3 * sizeof(double) * w->top);
  return BM_OK;
}


// src/world.js:771-803 (moved): the slots that moved in any step after `since` (and always in
// the last one), with bodies added since the last step: the awake list's in its order, then
// the log's from its end; a `since` older than the log reaches takes every body by its stamp.
int bm_world_set_parallel(bm_world* w, int32_t from) {
  if (!w) return BM_ERR_ARGUMENT;
  if (from < 1) return fail(w, BM_ERR_ARGUMENT, "bm_world_set_parallel: from must be 1 or more");
  w->parallel_from = from;
  return BM_OK;
}

int bm_world_set_broadphase(bm_world* w, int32_t mode) {
  if (!w) return BM_ERR_ARGUMENT;
  w->error.clear();
  if (mode != BM_BROAD_HOST && mode != BM_BROAD_DEVICE && mode != BM_BROAD_AUTO)
    return fail(w, BM_ERR_ARGUMENT, "bm_world_set_broadphase: mode is BM_BROAD_HOST, BM_BROAD_DEVICE or BM_BROAD_AUTO");
  if (mode == BM_BROAD_DEVICE && !w->owns_ctx)
    return fail(w, BM_ERR_ARGUMENT, "bm_world_set_broadphase: a batch's worlds take the host's broadphase");
  w->broad.mode = mode;
  return BM_OK;
}

int bm_world_events_of(const bm_world* w, bm_world_events* out) {
  if (!w || !out) return BM_ERR_ARGUMENT;
  const Events& e = w->events;
  out->count = e.count;
  out->type = e.type.data();
  out->a = e.a.data();
  out->b = e.b.data();
  out->point = e.point.data();
  out->normal = e.normal.data();
  out->speed = e.speed.data();
  return BM_OK;
}

int bm_world_moved(bm_world* w, int32_t since, int32_t* out, int32_t* count) {
  if (!w || !out || !count) return BM_ERR_ARGUMENT;
  w->error.clear();
  if (since < 0) return fail(w, BM_ERR_ARGUMENT, "moved: since must be a whole number, 0 or more");
  const int64_t after = std::min<int64_t>(since, (int64_t)w->step_count - 1);
  const int64_t* at = w->moved_at.data();
  int k = 0;
  if (after < w->log_floor) {
    for (int s = 0; s < w->top; s++)
      if (at[s] > after) out[k++] = s;
    *count = k;
    return BM_OK;
  }
  if (++w->seen_mark == 0) {
    std::fill(w->seen.begin(), w->seen.end(), 0u);
    w->seen_mark = 1;
  }
  const uint32_t mark = w->seen_mark;
  for (int i = 0; i < w->awake_count; i++) {
    const int s = w->awake_list[i];
    if (at[s] > after) {
      w->seen[s] = mark;
      out[k++] = s;
    }
  }
  for (int64_t i = (int64_t)w->log_slots.size() - 1; i >= 0 && w->log_at[i] > after; i--) {
    const int s = w->log_slots[i];
    if (s < w->top && at[s] > after && w->seen[s] != mark) {
      w->seen[s] = mark;
      out[k++] = s;
    }
  }
  *count = k;
  return BM_OK;
}

// ---- the adjoint: held trajectories of recorded worlds ----

int bm_world_record(bm_world* w, int32_t on) {
  if (!w) return BM_ERR_ARGUMENT;
  if (!on) {
    w->record.on = false;
    return BM_OK;
  }
  if (w->sleeping) return fail(w, BM_ERR_UNSUPPORTED, "record: a recorded world steps with sleeping off");
  record_start(w);
  return BM_OK;
}

int bm_world_recorded(const bm_world* w) { return w ? w->record.steps : 0; }

}  // extern "C"

struct bm_adjoint {
  bm_context* ctx = nullptr;
  std::string error;
  int bodies = 0, steps = 0;
  std::vector<double> state0, mass, friction, restitution;
  // The held trajectory as the device took it (bm_adjoint_held), and the arrays it points into.
  bm_held held{};
  std::vector<int32_t> i32[9];
  std::vector<uint8_t> u8[5];
  std::vector<uint32_t> u32;
  std::vector<double> f64[22];
};

namespace {
int adjoint_fail(bm_adjoint* a, int code, const std::string& m) {
  a->error = m;
  return code;
}
}  // namespace

extern "C" {

int bm_adjoint_create(bm_adjoint** out, int32_t device, bm_world* const* worlds, int32_t n, int32_t every) {
  if (!out || !worlds || n <= 0 || every <= 0) return BM_ERR_ARGUMENT;
  *out = nullptr;
  auto* A = new bm_adjoint();
  auto bad = [&](int code, const std::string& m) {
    A->error = m;
    *out = A;
    return code;
  };
  const int T = worlds[0] ? worlds[0]->record.steps : 0;
  for (int k = 0; k < n; k++) {
    const bm_world* w = worlds[k];
    if (!w) return bad(BM_ERR_ARGUMENT, "bm_adjoint_create: no world");
    const Record& R = w->record;
    const std::string which = "bm_adjoint_create: world " + std::to_string(k) + ": ";
    if (R.steps == 0) return bad(BM_ERR_ARGUMENT, which + "nothing recorded (bm_world_record, then steps)");
    if (R.steps != T) return bad(BM_ERR_ARGUMENT, which + "the worlds recorded different numbers of steps");
    if (!R.broken.empty()) return bad(BM_ERR_UNSUPPORTED, which + R.broken);
    if (w->catches != R.catches) return bad(BM_ERR_UNSUPPORTED, which + "a fast body was caught (the held step has no catch)");
    for (int s = 0; s < R.top; s++)
      if (w->lock[s] != 0 || w->joint_head[s] != 0)
        return bad(BM_ERR_UNSUPPORTED, which + "locked axes or joints");
  }
  // bodies
  std::vector<int32_t> body_offset(n + 1, 0), substeps(n);
  std::vector<double> dt(n), gravity(3 * n);
  for (int k = 0; k < n; k++) {
    body_offset[k + 1] = body_offset[k] + worlds[k]->record.top;
    substeps[k] = worlds[k]->substeps;
    dt[k] = worlds[k]->record.dt;
    for (int c = 0; c < 3; c++) gravity[3 * k + c] = worlds[k]->gravity[c];
  }
  const int nb = body_offset[n];
  std::vector<uint8_t> type(nb), shape(nb), tumbles(nb), capped(nb);
  std::vector<double> rounded_(nb), segment(nb), gscale(nb), damping(nb), adamping(nb), mass0(nb), inv_mass0(nb),
      inertia0(3 * nb), inv_inertia0(3 * nb), left(3 * nb), ceiling(nb), friction(nb), restitution(nb), state0;
  for (int k = 0; k < n; k++) {
    const bm_world* w = worlds[k];
    const Record& R = w->record;
    for (int s = 0; s < R.top; s++) {
      const int g = body_offset[k] + s;
      type[g] = w->type[s];
      shape[g] = w->shape[s];
      const double* I = &w->inertia[3 * s];
      tumbles[g] = w->type[s] == DYNAMIC && (I[0] != I[1] || I[1] != I[2]);
      capped[g] = w->shape[s] != SPHERE || w->framed[s] == 1;
      rounded_[g] = rounded(w, s);
      segment[g] = w->shape[s] == CAPSULE ? (w->dims[3 * s + 1] - 2 * w->dims[3 * s]) * 0.5 : 0;
      gscale[g] = w->gravity_scale[s];
      damping[g] = w->damping[s];
      adamping[g] = w->angular_damping[s];
      mass0[g] = w->mass[s];
      inv_mass0[g] = w->inv_mass[s];
      for (int c = 0; c < 3; c++) {
        inertia0[3 * g + c] = w->inertia[3 * s + c];
        inv_inertia0[3 * g + c] = w->inv_inertia[3 * s + c];
        left[3 * g + c] = R.left[3 * s + c];
      }
      ceiling[g] = R.ceiling[s];
      friction[g] = w->friction[s];
      restitution[g] = w->restitution[s];
    }
    state0.insert(state0.end(), R.initial.begin(), R.initial.end());
  }
  // rows: every world's first rows, then step by step, world by world
  std::vector<int32_t> first_base(n + 1, 0);
  for (int k = 0; k < n; k++) first_base[k + 1] = first_base[k] + (int)worlds[k]->record.first_count.size();
  const int first = first_base[n];
  std::vector<int32_t> row_offset(T * n + 1);
  {
    int at = first;
    for (int t = 0; t < T; t++)
      for (int k = 0; k < n; k++) {
        row_offset[t * n + k] = at;
        const Record& R = worlds[k]->record;
        at += R.row_offset[t + 1] - R.row_offset[t];
      }
    row_offset[T * n] = at;
  }
  const int rows = row_offset[T * n];
  std::vector<int32_t> a(rows, 0), b(rows, 0), src(rows, -1), starts((size_t)T * n * BM_STARTS), narrow(rows, 0);
  std::vector<uint8_t> count(rows, 0);
  std::vector<uint32_t> pid((size_t)P * rows, 0);
  std::vector<double> ln(3 * (size_t)rows, 0), ends(6 * P * (size_t)rows, 0), shift(3 * P * (size_t)rows, 0),
      tol(rows, 0), patch(rows, 0);
  std::vector<double> st_imp((size_t)P * first, 0), st_fr(3 * (size_t)first, 0), st_tw(3 * (size_t)first, 0),
      st_sl(3 * (size_t)first, 0);
  for (int k = 0; k < n; k++) {
    const Record& R = worlds[k]->record;
    for (int f = 0; f < (int)R.first_count.size(); f++) {
      const int g = first_base[k] + f;
      count[g] = R.first_count[f];
      for (int c = 0; c < P; c++) { pid[P * g + c] = R.first_pid[P * f + c]; st_imp[P * g + c] = R.first_impulse[P * f + c]; }
      for (int c = 0; c < 3; c++) {
        st_fr[3 * g + c] = R.first_friction[3 * f + c];
        st_tw[3 * g + c] = R.first_twist[3 * f + c];
        st_sl[3 * g + c] = R.first_slip[3 * f + c];
      }
    }
    // a world's own row numbers (steps one after another) onto the batch's
    std::vector<int32_t> global(R.a.size());
    for (int t = 0; t < T; t++)
      for (int f = R.row_offset[t]; f < R.row_offset[t + 1]; f++) global[f] = row_offset[t * n + k] + f - R.row_offset[t];
    for (int t = 0; t < T; t++) {
      for (int c = 0; c < BM_STARTS; c++) starts[((size_t)t * n + k) * BM_STARTS + c] = R.starts[(size_t)t * BM_STARTS + c];
      for (int f = R.row_offset[t]; f < R.row_offset[t + 1]; f++) {
        const int g = global[f];
        a[g] = R.a[f]; b[g] = R.b[f]; count[g] = R.count[f];
        narrow[g] = R.narrow[f];
        const int s = R.src[f];
        src[g] = s >= 0 ? global[s] : s <= -2 ? first_base[k] + (-2 - s) : -1;
        for (int c = 0; c < P; c++) pid[P * g + c] = R.pid[P * f + c];
        for (int c = 0; c < 3; c++) ln[3 * g + c] = R.local_normal[3 * f + c];
        for (int c = 0; c < 6 * P; c++) ends[6 * P * (size_t)g + c] = R.ends[6 * P * (size_t)f + c];
        for (int c = 0; c < 3 * P; c++) shift[3 * P * (size_t)g + c] = R.shift[3 * P * (size_t)f + c];
        tol[g] = R.tolerance[f];
        patch[g] = R.patch[f];
      }
    }
  }
  bm_held h{};
  h.worlds = n; h.steps = T; h.every = every; h.bodies = nb; h.rows = rows; h.first = first;
  h.body_offset = body_offset.data(); h.row_offset = row_offset.data(); h.starts = starts.data();
  h.narrow = narrow.data(); h.substeps = substeps.data(); h.dt = dt.data(); h.gravity = gravity.data();
  h.type = type.data(); h.shape = shape.data(); h.tumbles = tumbles.data(); h.capped = capped.data();
  h.rounded = rounded_.data(); h.segment = segment.data(); h.gravity_scale = gscale.data(); h.damping = damping.data();
  h.angular_damping = adamping.data(); h.mass0 = mass0.data(); h.inv_mass0 = inv_mass0.data();
  h.inertia0 = inertia0.data(); h.inv_inertia0 = inv_inertia0.data(); h.left = left.data(); h.ceiling = ceiling.data();
  h.a = a.data(); h.b = b.data(); h.src = src.data(); h.count = count.data(); h.pid = pid.data();
  h.local_normal = ln.data(); h.ends = ends.data(); h.shift = shift.data(); h.tolerance = tol.data();
  h.patch = patch.data();
  h.store_impulse = st_imp.data(); h.store_friction = st_fr.data(); h.store_twist = st_tw.data();
  h.store_slip = st_sl.data();
  if (int rc = bm_create(&A->ctx, device)) return bad(rc, "bm_adjoint_create: no device context");
  if (int rc = bm_held_load(A->ctx, &h)) return bad(rc, std::string("bm_adjoint_create: ") + bm_last_error(A->ctx));
  // Kept for bm_adjoint_held: a vector moved keeps its storage, so h's pointers stay good.
  {
    std::vector<int32_t>* I[9] = {&body_offset, &substeps, &row_offset, &a, &b, &src, &starts, &narrow, &first_base};
    std::vector<uint8_t>* U[5] = {&type, &shape, &tumbles, &capped, &count};
    std::vector<double>* D[22] = {&dt, &gravity, &rounded_, &segment, &gscale, &damping, &adamping, &mass0, &inv_mass0,
                                  &inertia0, &inv_inertia0, &left, &ceiling, &ln, &ends, &shift, &tol, &patch,
                                  &st_imp, &st_fr, &st_tw, &st_sl};
    for (int k = 0; k < 9; k++) A->i32[k] = std::move(*I[k]);
    for (int k = 0; k < 5; k++) A->u8[k] = std::move(*U[k]);
    for (int k = 0; k < 22; k++) A->f64[k] = std::move(*D[k]);
    A->u32 = std::move(pid);
    A->held = h;
  }
  A->bodies = nb;
  A->steps = T;
  A->state0 = std::move(state0);
  A->mass = A->f64[7];  // mass0, kept above
  A->friction = std::move(friction);
  A->restitution = std::move(restitution);
  *out = A;
  return BM_OK;
}

int bm_adjoint_destroy(bm_adjoint* a) {
  if (!a) return BM_OK;
  if (a->ctx) bm_destroy(a->ctx);
  delete a;
  return BM_OK;
}

const char* bm_adjoint_error(const bm_adjoint* a) { return a ? a->error.c_str() : "bm_adjoint: none"; }

int bm_adjoint_size(const bm_adjoint* a, int32_t* bodies, int32_t* steps) {
  if (!a) return BM_ERR_ARGUMENT;
  if (bodies) *bodies = a->bodies;
  if (steps) *steps = a->steps;
  return BM_OK;
}

int bm_adjoint_recorded(const bm_adjoint* a, double* state0, double* mass, double* friction, double* restitution) {
  if (!a || !a->ctx) return BM_ERR_ARGUMENT;
  if (state0) std::memcpy(state0, a->state0.data(), a->state0.size() * sizeof(double));
  if (mass) std::memcpy(mass, a->mass.data(), a->mass.size() * sizeof(double));
  if (friction) std::memcpy(friction, a->friction.data(), a->friction.size() * sizeof(double));
  if (restitution) std::memcpy(restitution, a->restitution.data(), a->restitution.size() * sizeof(double));
  return BM_OK;
}

int bm_adjoint_forward(bm_adjoint* a, const double* state0, const double* mass, const double* friction,
                       const double* restitution, double* state_out) {
  if (!a || !a->ctx || !state0 || !mass || !friction || !restitution || !state_out) return BM_ERR_ARGUMENT;
  a->error.clear();
  if (int rc = bm_held_forward(a->ctx, state0, mass, friction, restitution, state_out))
    return adjoint_fail(a, rc, std::string("forward: ") + bm_last_error(a->ctx));
  return BM_OK;
}

int bm_adjoint_backward(bm_adjoint* a, const double* state_bar, double* state0_bar, double* mass_bar,
                        double* friction_bar, double* restitution_bar) {
  if (!a || !a->ctx || !state_bar || !state0_bar || !mass_bar || !friction_bar || !restitution_bar)
    return BM_ERR_ARGUMENT;
  a->error.clear();
  if (int rc = bm_held_backward(a->ctx, state_bar, state0_bar, mass_bar, friction_bar, restitution_bar))
    return adjoint_fail(a, rc, std::string("backward: ") + bm_last_error(a->ctx));
  return BM_OK;
}

int bm_adjoint_times(const bm_adjoint* a, double* forward_ms, double* backward_ms) {
  if (!a || !a->ctx) return BM_ERR_ARGUMENT;
  return bm_held_times(a->ctx, forward_ms, backward_ms);
}

int bm_adjoint_held(const bm_adjoint* a, bm_held* out) {
  if (!a || !a->ctx || !out) return BM_ERR_ARGUMENT;
  *out = a->held;
  return BM_OK;
}

int bm_adjoint_branches(bm_adjoint* a, uint64_t* per_world) {
  if (!a || !a->ctx || !per_world) return BM_ERR_ARGUMENT;
  a->error.clear();
  if (int rc = bm_held_branches(a->ctx, per_world)) return adjoint_fail(a, rc, std::string("branches: ") + bm_last_error(a->ctx));
  return BM_OK;
}

}  // extern "C"
