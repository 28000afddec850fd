// The broadphase's data-parallel passes on the device: each awake body's bounds and fat-box
// test (src/world.js:1574-1630, src/broadphase.js:140-159), each marked body's walk of the two
// trees (src/broadphase.js:161-179 candidates, src/tree.js:259-275 query) and each pair's
// fat-box overlap (src/pairs.js:340-356 holds). The order-defining parts (re-filing a body in
// the tree, handing the walks' leaves to the pair list, removing pairs) stay with the host,
// which takes these results in the reference's order.
#include "broad.h"

namespace bm {

// One awake body a lane. The sixteen results go out field-major, so a wave's stores of one
// field are one contiguous run.
__global__ void __launch_bounds__(256) bounds_kernel(BoundsDev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int s = d.list[i];
  double o[16];
  broad::bounds(d.bodies, s, d.dt, d.gl, o);
  for (int k = 0; k < 16; k++) d.out[(size_t)k * d.n + i] = o[k];
  uint8_t r = 0;
  if (d.which[s] == broad::MOVING_TREE) r = broad::contains(d.moving.node[d.leaf[s]].box, o + 10) ? 0 : 1;
  d.refile[i] = r;
}

// One marked body a lane, a wave a workgroup: the walks once, keeping the first FOUND leaves
// met and counting them all, a reservation of that many places, the kept leaves written there,
// and only for a body that met more the walks again to write the rest, so each body's leaves
// stay in walk order. Each lane's stack and kept leaves live in LDS, entry k of lane l at
// k * 64 + l.
constexpr int FOUND = 32;
__global__ void __launch_bounds__(64) query_kernel(QueryDev d) {
  __shared__ int stacks[broad::STACK * 64];
  __shared__ int kept[FOUND * 64];
  int* stack = stacks + threadIdx.x;
  int* mine = kept + threadIdx.x;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int a = d.moved[i];
  const int t = d.which[a];
  d.counts[i] = 0;
  d.starts[i] = 0;
  if (t < 0) return;
  double q[6];
  const double* lb = d.trees[t].node[d.leaf[a]].box;
  for (int k = 0; k < 6; k++) q[k] = lb[k];
  const bool stat = d.type[a] == broad::STATIC;
  int c = 0;
  auto count = [&](int b) {
    if (b == a) return;
    if (c < FOUND) mine[c * 64] = b;
    c++;
  };
  bool ok = true;
  if (!stat) ok = broad::query(d.trees[broad::STATIC_TREE], q, stack, 64, count);
  if (ok) ok = broad::query(d.trees[broad::MOVING_TREE], q, stack, 64, count);
  if (!ok) {
    atomicOr(d.flags, 1);
    return;
  }
  const int base = atomicAdd(d.total, c);
  d.counts[i] = c;
  d.starts[i] = base;
  if (c == 0 || base + c > d.cap) return;
  for (int k = 0; k < c && k < FOUND; k++) d.hits[base + k] = mine[k * 64];
  if (c <= FOUND) return;
  // more than were kept: the same walks again, writing from the first leaf not kept on
  int seen = 0;
  auto put = [&](int b) {
    if (b == a) return;
    if (seen >= FOUND && seen < c) d.hits[base + seen] = b;
    seen++;
  };
  if (!stat) broad::query(d.trees[broad::STATIC_TREE], q, stack, 64, put);
  broad::query(d.trees[broad::MOVING_TREE], q, stack, 64, put);
}

__global__ void __launch_bounds__(256) holds_kernel(HoldsDev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int a = d.a[i], b = d.b[i];
  // Every pair's bodies are filed in a tree; a pair that names one that is not stays held.
  if (d.which[a] < 0 || d.which[b] < 0) {
    d.apart[i] = 0;
    return;
  }
  const double* fa = d.trees[d.which[a]].node[d.leaf[a]].box;
  const double* fb = d.trees[d.which[b]].node[d.leaf[b]].box;
  d.apart[i] = broad::overlap(fa, fb) ? 0 : 1;
}

__global__ void __launch_bounds__(256) scatter_kernel(int n, const int* slots, const double* rows, double* position,
                                                      double* rotation, double* velocity, double* angular) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t s = slots[i];
  const double* r = rows + 13 * (size_t)i;
  for (int k = 0; k < 3; k++) position[3 * s + k] = r[k];
  for (int k = 0; k < 4; k++) rotation[4 * s + k] = r[3 + k];
  for (int k = 0; k < 3; k++) velocity[3 * s + k] = r[7 + k];
  for (int k = 0; k < 3; k++) angular[3 * s + k] = r[10 + k];
}

void launch_broad_scatter(int n, const int* slots, const double* rows, double* position, double* rotation,
                          double* velocity, double* angular, hipStream_t stream) {
  if (n > 0) scatter_kernel<<<(n + 255) / 256, 256, 0, stream>>>(n, slots, rows, position, rotation, velocity, angular);
}

void launch_bounds(const BoundsDev& d, hipStream_t stream) {
  if (d.n > 0) bounds_kernel<<<(d.n + 255) / 256, 256, 0, stream>>>(d);
}

void launch_query(const QueryDev& d, hipStream_t stream) {
  if (d.n > 0) query_kernel<<<(d.n + 63) / 64, 64, 0, stream>>>(d);
}

void launch_holds(const HoldsDev& d, hipStream_t stream) {
  if (d.n > 0) holds_kernel<<<(d.n + 255) / 256, 256, 0, stream>>>(d);
}

}  // namespace bm
