// Constraint-graph colouring on the MI355X (src/graph.js:105-214), one workgroup per world.
//
// The reference colours manifolds greedily in the order narrow() found them: each takes
// the lowest of 32 colours that no earlier manifold on one of its moving bodies holds
// (or the overflow when all 32 are held), so a manifold's colour depends only on the
// colours of the earlier manifolds it shares a moving body with. That makes it a
// priority fixed point: a manifold whose earlier neighbours all have their colours can
// take its own, the same one the serial pass gives it. Each round every undecided
// manifold checks its earlier neighbours (through a per-body list built with LDS-free
// atomics in HBM scratch) and decides if it can; rounds run until none is left, as
// many as the longest chain of dependent manifolds, not as many as manifolds.
//
// A world whose data fits in LDS is coloured the reference's own way instead, one manifold
// after another by one lane, with the held masks, the ends and the colours in LDS; every
// step of that loop is an LDS access, so a long chain of dependent manifolds costs an LDS
// round trip each rather than a pass of the whole workgroup through HBM.
//
// The rows are then laid out as the reference's stable counting sorts leave them: by
// colour, then by b's slot, then by the order found. That key (colour, b, m) is unique,
// so a bitonic sort of it in scratch gives exactly that order. Resting rows (between
// frozen or static bodies, src/graph.js RESTING) go last and get rest()'s values
// (src/contact.js:1318-1329). Then every column narrow() writes is moved to its row
// (src/graph.js:178-214), through spare buffers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/device/device_radix_sort.hpp>

#include "colour.h"

namespace bm {

namespace {
constexpr int DYNAMIC = 0, STATIC = 1;
constexpr int COLOURS = 32, OVERFLOW = 32, RESTING = 33;
constexpr int UNDECIDED = -1;
constexpr int STREAM_OWNERS = 4096;  // the streamed walk's owner slots (a power of two)

__device__ __forceinline__ int lowest_free(int32_t held) {
  const int32_t bit = ~held & (held + 1);
  return 31 - __clz((uint32_t)bit);
}
}  // namespace

// One workgroup a world. Scratch per world (at the world's manifold / body offsets):
// deg[nb + 1] (per body), adj[2 * count], keys[pow2 >= count] (64-bit), state[count].
__global__ void __launch_bounds__(256) colour_kernel(ColourDev d) {
  __shared__ int flag;
  __shared__ int runs[RESTING + 2];
  extern __shared__ __align__(16) unsigned char lds[];
  for (int at = blockIdx.x; at < d.n_worlds; at += gridDim.x) {
    const int wld = d.world_order ? d.world_order[at] : at;
    if (d.in_lds) {
      const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
      const int nb = d.body_offset[wld + 1] - boff, n = d.manifold_offset[wld + 1] - moff;
      const int width = d.key_offset[wld + 1] - d.key_offset[wld];
      // The keys take the place of held, ends and moving once the colours are known.
      uint64_t* keys = reinterpret_cast<uint64_t*>(lds);
      int32_t* held = reinterpret_cast<int32_t*>(lds);
      int2* ends = reinterpret_cast<int2*>(held + nb + (nb & 1));
      uint8_t* moving = reinterpret_cast<uint8_t*>(ends + n);
      int* col = reinterpret_cast<int*>(lds + colour_lds_front(nb, n, width));
      const int* A = d.a + moff;
      const int* B = d.b + moff;
      const uint8_t* type = d.type + boff;
      const uint8_t* frozen = d.frozen + boff;
      __syncthreads();
      for (int s = threadIdx.x; s < nb; s += blockDim.x) {
        held[s] = 0;
        moving[s] = type[s] == DYNAMIC && frozen[s] == 0;
      }
      for (int k = threadIdx.x; k < RESTING + 2; k += blockDim.x) runs[k] = 0;
      __syncthreads();
      for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int a = A[m], b = B[m];
        ends[m] = make_int2(a, b);
        // Resting rows: between frozen or static bodies (src/graph.js:114-124).
        const bool rest_a = frozen[a] != 0 || type[a] == STATIC, rest_b = frozen[b] != 0 || type[b] == STATIC;
        col[m] = rest_a && rest_b ? RESTING : UNDECIDED;
      }
      __syncthreads();
      // src/graph.js:125-140, in order, by the first wave: 64 manifolds at a time, each
      // lane fetching one's ends and flags, then the wave walking them one by one with the
      // ends read out of lane registers, so each step waits only on its two held masks.
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int base = 0; base < n; base += 64) {
          const int m = base + lane;
          int ea = 0, eb = 0, flags = 0;
          if (m < n) {
            const int2 e = ends[m];
            ea = e.x;
            eb = e.y;
            flags = (moving[e.x] ? 1 : 0) | (moving[e.y] ? 2 : 0) | (col[m] == RESTING ? 4 : 0);
          }
          int mine = RESTING;
          const int count = n - base < 64 ? n - base : 64;
          for (int j = 0; j < count; j++) {
            const int f = __builtin_amdgcn_readlane(flags, j);
            if (f & 4) continue;
            const int a = __builtin_amdgcn_readlane(ea, j), b = __builtin_amdgcn_readlane(eb, j);
            const int32_t h = ((f & 1) ? held[a] : 0) | ((f & 2) ? held[b] : 0);
            int c;
            if (h == -1) {
              c = OVERFLOW;
            } else {
              const int32_t bit = ~h & (h + 1);
              c = 31 - __clz((uint32_t)bit);
              if (lane == 0) {
                if (f & 1) held[a] |= bit;
                if (f & 2) held[b] |= bit;
              }
            }
            if (lane == j) mine = c;
          }
          if (m < n) col[m] = mine;
        }
      }
      __syncthreads();
      int* colour = d.colour + moff;
      for (int i = threadIdx.x; i < width; i += blockDim.x) {
        if (i < n) {
          const int c = col[i];
          colour[i] = c;
          keys[i] = ((uint64_t)c << 56) | ((uint64_t)(uint32_t)B[i] << 32) | (uint32_t)i;
          atomicAdd(&runs[c + 1], 1);
        } else {
          keys[i] = ~0ull;
        }
      }
      __syncthreads();
      for (int size = 2; size <= width; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
          for (int i = threadIdx.x; i < width; i += blockDim.x) {
            const int j = i ^ stride;
            if (j > i) {
              const uint64_t x = keys[i], y = keys[j];
              const bool up = (i & size) == 0;
              if ((x > y) == up) { keys[i] = y; keys[j] = x; }
            }
          }
          __syncthreads();
        }
      }
      int* order = d.order + moff;
      int* row = d.row + moff;
      for (int o = threadIdx.x; o < n; o += blockDim.x) {
        const int m = (int)(uint32_t)(keys[o] & 0xffffffffu);
        order[o] = m;
        row[m] = o;
      }
      if (threadIdx.x == 0) {
        int* starts = d.starts + BM_STARTS_DEV * wld;
        starts[0] = 0;
        for (int k = 0; k <= RESTING; k++) starts[k + 1] = starts[k] + runs[k + 1];
        d.solved[wld] = starts[RESTING];
      }
      __syncthreads();
      continue;
    }
    const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
    const int nb = d.body_offset[wld + 1] - boff, n = d.manifold_offset[wld + 1] - moff;
    const int* A = d.a + moff;
    const int* B = d.b + moff;
    int* colour = d.colour + moff;
    int* deg = d.deg + boff + wld;           // nb + 1 entries a world
    int* adj = d.adj + 2 * moff;             // 2 per manifold
    uint64_t* keys = d.keys + d.key_offset[wld];
    const int width = d.key_offset[wld + 1] - d.key_offset[wld];
    const uint8_t* type = d.type + boff;
    const uint8_t* frozen = d.frozen + boff;
    __syncthreads();
    for (int s = threadIdx.x; s <= nb; s += blockDim.x) deg[s] = 0;
    for (int k = threadIdx.x; k < RESTING + 2; k += blockDim.x) runs[k] = 0;
    __syncthreads();
    // Which ends move, and the resting rows (src/graph.js:114-124).
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
      const int a = A[m], b = B[m];
      const bool fa = frozen[a] != 0, fb = frozen[b] != 0;
      if ((fa || type[a] == STATIC) && (fb || type[b] == STATIC)) {
        colour[m] = RESTING;
        continue;
      }
      colour[m] = UNDECIDED;
      if (type[a] == DYNAMIC && !fa) atomicAdd(&deg[a + 1], 1);
      if (type[b] == DYNAMIC && !fb) atomicAdd(&deg[b + 1], 1);
    }
    __syncthreads();
    // Body lists' starts: an inclusive scan of deg[1..nb], one wave strided.
    if (threadIdx.x == 0) {
      for (int s = 1; s <= nb; s++) deg[s] += deg[s - 1];
    }
    __syncthreads();
    // Fill: deg[s] walks up to the next body's start.
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
      if (colour[m] == RESTING) continue;
      const int a = A[m], b = B[m];
      if (type[a] == DYNAMIC && frozen[a] == 0) adj[atomicAdd(&deg[a], 1)] = m;
      if (type[b] == DYNAMIC && frozen[b] == 0) adj[atomicAdd(&deg[b], 1)] = m;
    }
    __syncthreads();
    // deg[s] is now body s's end, body s-1's end its start (deg[-1] read as 0).
    do {
      __syncthreads();
      if (threadIdx.x == 0) flag = 0;
      __syncthreads();
      for (int m = threadIdx.x; m < n; m += blockDim.x) {
        if (__hip_atomic_load(&colour[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != UNDECIDED) continue;
        const int a = A[m], b = B[m];
        int32_t held = 0;
        bool ready = true;
        for (int end = 0; end < 2 && ready; end++) {
          const int s = end == 0 ? a : b;
          if (type[s] != DYNAMIC || frozen[s] != 0) continue;
          const int from = s == 0 ? 0 : deg[s - 1], to = deg[s];
          for (int k = from; k < to; k++) {
            const int o = adj[k];
            if (o >= m) continue;
            const int c = __hip_atomic_load(&colour[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (c == UNDECIDED) { ready = false; break; }
            if (c < COLOURS) held |= (int32_t)(1u << c);
          }
        }
        if (!ready) { flag = 1; continue; }
        __hip_atomic_store(&colour[m], held == -1 ? OVERFLOW : lowest_free(held), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      __syncthreads();
    } while (flag);
    // Keys (colour, b, m), padded past n with the largest key; runs counted per colour.
    for (int i = threadIdx.x; i < width; i += blockDim.x) {
      if (i < n) {
        const int c = colour[i];
        keys[i] = ((uint64_t)c << 56) | ((uint64_t)(uint32_t)B[i] << 32) | (uint32_t)i;
        atomicAdd(&runs[c + 1], 1);
      } else {
        keys[i] = ~0ull;
      }
    }
    __syncthreads();
    // Bitonic sort of `width` (a power of two) keys.
    for (int size = 2; size <= width; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = threadIdx.x; i < width; i += blockDim.x) {
          const int j = i ^ stride;
          if (j > i) {
            const uint64_t x = keys[i], y = keys[j];
            const bool up = (i & size) == 0;
            if ((x > y) == up) { keys[i] = y; keys[j] = x; }
          }
        }
        __syncthreads();
      }
    }
    int* order = d.order + moff;
    int* row = d.row + moff;
    for (int o = threadIdx.x; o < n; o += blockDim.x) {
      const int m = (int)(uint32_t)(keys[o] & 0xffffffffu);
      order[o] = m;
      row[m] = o;
    }
    if (threadIdx.x == 0) {
      int* starts = d.starts + BM_STARTS_DEV * wld;
      starts[0] = 0;
      for (int k = 0; k <= RESTING; k++) starts[k + 1] = starts[k] + runs[k + 1];
      d.solved[wld] = starts[RESTING];
    }
    __syncthreads();
  }
}

// A single world's manifolds packed for the streamed walk below, one 8-byte word each: a's
// slot, then b's slot with the flags in its top byte: 1 a moves, 2 b moves, 4 resting (both
// ends frozen or static, src/graph.js:114-124). Done across the grid, so the walk's fetches
// are plain loads with nothing that depends on them.
__global__ void stream_ends_kernel(ColourDev d, int wld, int2* ends) {
  const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
  const int n = d.manifold_offset[wld + 1] - moff;
  const uint8_t* type = d.type + boff;
  const uint8_t* frozen = d.frozen + boff;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
    const int a = d.a[moff + m], b = d.b[moff + m];
    const bool move_a = type[a] == DYNAMIC && frozen[a] == 0, move_b = type[b] == DYNAMIC && frozen[b] == 0;
    const bool rest = (frozen[a] != 0 || type[a] == STATIC) && (frozen[b] != 0 || type[b] == STATIC);
    const int flags = (move_a ? 1 : 0) | (move_b ? 2 : 0) | (rest ? 4 : 0);
    ends[m] = make_int2(a, b | (flags << 24));
  }
}

// One world too large to colour in LDS whose bodies' held masks fit there: the reference's
// walk (src/graph.js:125-140) by the first wave, the masks in LDS, the packed manifolds
// streamed from HBM 64 at a time (each lane fetching its next one while the wave colours the
// current 64). Within the 64 the colours are the serial walk's, found in rounds: a manifold
// whose earlier undecided neighbours among the 64 (those sharing a moving body with it) are
// all decided reads its masks, takes the lowest colour they leave free and ORs it in; two
// manifolds that decide in one round share no moving body, so each reads exactly the colours
// of the manifolds before it on its bodies, as the serial walk does. A manifold is ready when
// it is the lowest undecided lane holding each of its moving bodies' owner slots (a table
// of STREAM_OWNERS, by body number modulo its size: two bodies on one slot only make a lane
// wait a round longer, and the lowest undecided lane is always ready). The rounds are as
// many as the longest chain of dependent manifolds in the 64. Then the runs per colour,
// starts and solved, and the sort keys (colour, b) with the manifold numbers, for a stable
// radix sort that leaves the rows as the reference's counting sorts do.
__global__ void __launch_bounds__(256) colour_stream_kernel(ColourDev d, int wld, const int2* ends, uint64_t* keys,
                                                            int* values) {
  __shared__ int runs[RESTING + 2];
  extern __shared__ __align__(16) unsigned char lds[];
  const int boff = d.body_offset[wld], moff = d.manifold_offset[wld];
  const int nb = d.body_offset[wld + 1] - boff, n = d.manifold_offset[wld + 1] - moff;
  int* stage = reinterpret_cast<int*>(lds);
  int* owner = stage + 128;
  int32_t* held = owner + STREAM_OWNERS;
  const int* B = d.b + moff;
  int* colour = d.colour + moff;
  for (int s = threadIdx.x; s < nb; s += blockDim.x) held[s] = 0;
  for (int k = threadIdx.x; k < RESTING + 2; k += blockDim.x) runs[k] = 0;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    // Each lane's next manifold waits in LDS, in its own slot (a compiler barrier after each
    // store keeps it there rather than in registers): its load is issued before the rounds
    // and waited for after them.
    int* slot = &stage[2 * lane];
    const int2 first = lane < n ? ends[lane] : make_int2(0, 0);
    slot[0] = first.x;
    slot[1] = first.y;
    for (int base = 0; base < n; base += 64) {
      const int2 cur = make_int2(slot[0], slot[1]);
      const int2 next = base + 64 + lane < n ? ends[base + 64 + lane] : make_int2(0, 0);
      const int ea = cur.x, eb = cur.y & 0xffffff, flags = cur.y >> 24;
      const bool move_a = (flags & 1) != 0, move_b = (flags & 2) != 0;
      int* own_a = &owner[ea & (STREAM_OWNERS - 1)];
      int* own_b = &owner[eb & (STREAM_OWNERS - 1)];
      bool undecided = base + lane < n;
      int mine = RESTING;
      while (__any(undecided)) {
        // The owner slots: cleared by every undecided lane on them, then each takes the lowest
        // such lane (the wave's LDS operations are done in the order issued; the compiler
        // barriers keep that order in the code).
        if (undecided && move_a) *own_a = 64;
        if (undecided && move_b) *own_b = 64;
        __asm__ __volatile__("" ::: "memory");
        if (undecided && move_a) atomicMin(own_a, lane);
        if (undecided && move_b) atomicMin(own_b, lane);
        __asm__ __volatile__("" ::: "memory");
        const bool ready = undecided && (!move_a || *own_a == lane) && (!move_b || *own_b == lane);
        if (ready) {
          const uint32_t h = (uint32_t)((move_a ? held[ea] : 0) | (move_b ? held[eb] : 0));
          if (flags & 4) {
            mine = RESTING;  // no colour taken
          } else if (h == 0xffffffffu) {
            mine = OVERFLOW;  // every colour held: none taken
          } else {
            const uint32_t bit = ~h & (h + 1u);
            mine = 31 - __clz(bit);
            if (move_a) held[ea] |= (int32_t)bit;
            if (move_b) held[eb] |= (int32_t)bit;
          }
          undecided = false;
        }
        __asm__ __volatile__("" ::: "memory");
      }
      if (base + lane < n) colour[base + lane] = mine;
      slot[0] = next.x;
      slot[1] = next.y;
      __asm__ __volatile__("" ::: "memory");
    }
  }
  __syncthreads();
  for (int m = threadIdx.x; m < n; m += blockDim.x) {
    const int c = colour[m];
    atomicAdd(&runs[c + 1], 1);
    keys[m] = ((uint64_t)c << 32) | (uint32_t)B[m];
    values[m] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int* starts = d.starts + BM_STARTS_DEV * wld;
    starts[0] = 0;
    for (int k = 0; k <= RESTING; k++) starts[k + 1] = starts[k] + runs[k + 1];
    d.solved[wld] = starts[RESTING];
  }
}

// order[o] = the o-th manifold in the sorted order, row[m] its row.
__global__ void order_kernel(const int* sorted, int n, int* order, int* row) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int m = sorted[o];
    order[o] = m;
    row[m] = o;
  }
}

size_t colour_stream_lds(int bodies) { return ((size_t)(128 + STREAM_OWNERS + bodies) * 4 + 15) & ~(size_t)15; }

size_t colour_stream_scratch(int manifolds) {
  size_t temp = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (const int*)nullptr, (int*)nullptr, manifolds, 0, 38);
  return temp;
}

int launch_colour_stream(const ColourDev& d, int wld, int manifolds, int bodies, void* scratch, size_t scratch_bytes,
                         hipStream_t stream) {
  // Scratch: keys in and out, values in and out, then the sort's own; the packed manifolds
  // take the place of the keys out until the sort writes them.
  uint64_t* keys = static_cast<uint64_t*>(scratch);
  uint64_t* keys_out = keys + manifolds;
  int* values = reinterpret_cast<int*>(keys_out + manifolds);
  int* values_out = values + manifolds;
  char* temp = reinterpret_cast<char*>(values_out + manifolds + (manifolds & 1));
  const size_t used = (size_t)(temp - static_cast<char*>(scratch));
  size_t temp_bytes = scratch_bytes > used ? scratch_bytes - used : 0;
  int2* ends = reinterpret_cast<int2*>(keys_out);
  const int blocks = (manifolds + 255) / 256 < 1024 ? (manifolds + 255) / 256 : 1024;
  hipLaunchKernelGGL(stream_ends_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, stream, d, wld, ends);
  hipLaunchKernelGGL(colour_stream_kernel, dim3(1), dim3(256), colour_stream_lds(bodies), stream, d, wld, ends, keys,
                     values);
  // Colour (6 bits) above b's slot (32 bits): the stable sort keeps the order found within a
  // colour and slot.
  if (hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_out, values, values_out, manifolds, 0, 38,
                                         stream) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(order_kernel, dim3((manifolds + 255) / 256), dim3(256), 0, stream, values_out, manifolds,
                     d.order, d.row);
  return 0;
}

// Moves column `c` of every world's manifolds to their rows: to[o] = from[order[o]].
__global__ void arrange_kernel(ArrangeDev d) {
  const int col = blockIdx.y;
  const ArrangeColumn c = d.columns[col];
  const long total = (long)d.manifolds * c.per;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long o = i / c.per, k = i - o * c.per;
    const int w = d.world_of[o];
    const long m = d.manifold_offset[w] + d.order[o];
    if (c.bytes == 8) static_cast<uint64_t*>(c.to)[i] = static_cast<const uint64_t*>(c.from)[m * c.per + k];
    else if (c.bytes == 4) static_cast<uint32_t*>(c.to)[i] = static_cast<const uint32_t*>(c.from)[m * c.per + k];
    else static_cast<uint8_t*>(c.to)[i] = static_cast<const uint8_t*>(c.from)[m * c.per + k];
  }
}

// rest() (src/contact.js:1318-1329) on each world's resting rows, after they are laid out.
__global__ void rest_kernel(RestDev d) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.manifolds) return;
  const int w = d.world_of[m];
  if (m - d.manifold_offset[w] < d.solved[w]) return;
  double least = __builtin_inf();
  for (int k = 0; k < d.points[m]; k++) {
    const long i = 4 * m + k;
    if (d.separation[i] < least) least = d.separation[i];
    d.max_normal_impulse[i] = d.normal_impulse[i];
    d.relative_velocity[i] = 0;
  }
  d.gap[m] = least;
}

// Each manifold's world, for the flat kernels above.
__global__ void world_of_kernel(const int* manifold_offset, int n_worlds, int* world_of) {
  const int w = blockIdx.x;
  if (w >= n_worlds) return;
  for (int m = manifold_offset[w] + threadIdx.x; m < manifold_offset[w + 1]; m += blockDim.x) world_of[m] = w;
}

void launch_colour(const ColourDev& d, size_t lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(colour_kernel, dim3(d.n_worlds), dim3(256), d.in_lds ? lds_bytes : 0, stream, d);
}

size_t colour_lds_limit(int device) {
  int most = 0;
  if (hipDeviceGetAttribute(&most, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return 0;
  const size_t own = (RESTING + 2 + 1) * sizeof(int) + 64;
  return (size_t)most > own ? (size_t)most - own : 0;
}

void launch_world_of(const int* manifold_offset, int n_worlds, int* world_of, hipStream_t stream) {
  hipLaunchKernelGGL(world_of_kernel, dim3(n_worlds), dim3(256), 0, stream, manifold_offset, n_worlds, world_of);
}

void launch_arrange(const ArrangeDev& d, int columns, hipStream_t stream) {
  if (d.manifolds == 0) return;
  hipLaunchKernelGGL(arrange_kernel, dim3(256, columns), dim3(256), 0, stream, d);
}

void launch_rest(const RestDev& d, hipStream_t stream) {
  if (d.manifolds == 0) return;
  hipLaunchKernelGGL(rest_kernel, dim3((d.manifolds + 255) / 256), dim3(256), 0, stream, d);
}

}  // namespace bm
