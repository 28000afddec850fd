// Device-side view of a batch of shape pairs for the narrowphase (collide.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {

// The hulls shape records name by number: each hull's tables in its body's frame
// (src/hull.js Hull), with the reference's unit cube (src/collide.js CUBE) among them.
constexpr int HULL_MAX_FACES = 64, HULL_MAX_EDGES = 96, HULL_MAX_CORNERS = 16;
struct HullTable {
  int n, cube;                 // hulls; the unit cube's number
  const int* info;             // 8 a hull: vertex, plane, start, loop and edge offsets; V, F, E
  const double* dinfo;         // 2 a hull: inner, reach
  const double *vertices, *planes;
  const int *start, *loops, *edges, *sides;  // start: F + 1 a hull, counted in the hull's loops
};

struct CollideDev {
  int n;
  const double *a, *b;             // shape records: kind, centre 3, axes 9, half 3 (a hull: its number), weight 3
  const double *margin, *tolerance;
  int* count;                      // -1 for a pair of kinds not handled (capsules, hulls)
  double *normal, *points, *separations;
  uint32_t* ids;
  double* loss;                    // Manifold.loss: how far a hull face falls short of the best axis
  HullTable hulls;
};

void launch_collide(const CollideDev& d, hipStream_t stream);

struct MeetDev {
  int n;
  const double *a, *b;  // body records: the shape record (19), rotation 4, velocity 3, spin 3,
                        // gravity scale, reach, flags (1 dynamic, 2 spin capped)
  const int *ia, *ib;   // when set, pair k's bodies are records ia[k] of a and ib[k] of b (one
                        // table of each body once); else records k
  const double* pair;   // margin, tolerance, dt, gravity 3
  int *count, *flags;   // flags: 1 shifted, 2 a time of impact was asked
  double *normal, *points, *separations, *shift;
  uint32_t* ids;
  double* loss;
  HullTable hulls;
};

void launch_meet(const MeetDev& d, hipStream_t stream);
// A meet pass's results packed by pair: row[k] the packed row of pair k when it found a
// manifold (count[k] > 0, rows in pair order), else -1; *found the rows; the packed columns
// have room for n rows; sums: room for an int per 256 pairs (the workgroups' counts, then
// their first rows). Workgroups count, one workgroup numbers them, then a lane a pair copies
// its row.
struct MeetPacked {
  int *row, *found, *sums;
  double *normal, *points, *separations, *shift, *loss;
  uint32_t* ids;
};
void launch_meet_pack(const MeetDev& d, const MeetPacked& p, hipStream_t stream);

// Heightfields (src/mesh.js:288 Heightfield) by number: field f has info[2f] columns and
// info[2f + 1] rows of heights from heights + off[f], spacing dinfo[3f], x0 dinfo[3f + 1],
// z0 dinfo[3f + 2] (its lowest corner in its body's frame).
struct FieldTable {
  int n;
  const int* info;
  const long long* off;
  const double* dinfo;
  const double* heights;
};

// touchLevel (src/triangle.js:1078) for n shapes, one a lane: shape record k against field
// field[k] whose body is at position 3k turned by axes 9k, looking in box 6k (the field's
// frame). scratch (n + 1 prefix): how many points lane k may find, 7 a triangle under its
// box; sd / su its working storage, 9 doubles and 3 words a point. Out: manifolds[k] (every
// direction found; slots past cap are not written), then cap manifold slots a lane.
struct LevelDev {
  int n, cap;
  const int* field;
  const double *position, *axes, *shapes, *boxes, *margin, *tolerance;
  const long long* scratch;
  double* sd;
  uint32_t* su;
  int *manifolds, *count;
  double *normal, *points, *separations;
  uint32_t* ids;
  HullTable hulls;
  FieldTable fields;
};

void launch_level(const LevelDev& d, hipStream_t stream);

// The no-allocation form of the level pass: its working memory is a fixed block of device
// memory (a file-scope __device__ array of collide.hip), one copy on each device, holding for
// up to LEVEL_NOALLOC_PAIRS pairs a call each pair's offset into the found points, a status
// block, and up to LEVEL_NOALLOC_POINTS found points (9 doubles and 3 words each).
constexpr int LEVEL_NOALLOC_PAIRS = 1 << 16;
constexpr long long LEVEL_NOALLOC_POINTS = 1 << 18;
// The block's address on the current device (taken with hipGetSymbolAddress the first time
// a device asks, then kept for that device); hipSuccess or the error that call returned.
hipError_t level_noalloc_block(int device, void** out);
// Status the plan leaves in the block: [0] points the pairs need (7 a triangle under each box),
// [1] flags (1 a shape kind other than 0-3, 2 a hull number the table lacks, 4 a field number
// the table lacks, 8 a pair needing more points than the block holds), [2] the first pair so
// flagged (n if none).
constexpr int LEVEL_PLAN_KIND = 1, LEVEL_PLAN_HULL = 2, LEVEL_PLAN_FIELD = 4, LEVEL_PLAN_POINTS = 8;
// Checks every pair's shape kind, hull and field numbers, and gives each pair its share of the
// block's found points (one workgroup); then the pass itself may be launched on the block.
void launch_level_plan(const LevelDev& d, void* block, hipStream_t stream);
// Where in the block the plan leaves its status (3 words).
const void* level_plan_status(const void* block);
// The LevelDev of a planned pass: offsets and working storage in the block.
void level_on_block(LevelDev& d, void* block);
// The pass's results packed by pair: start[k] (n + 1 of them) the first packed row of pair k,
// each pair's min(manifolds[k], cap) slots in order; out_* the packed columns (room for n * cap
// rows each). One workgroup scans, then a lane a pair copies its slots.
struct LevelPacked {
  int* start;
  int* count;
  double *normal, *points, *separations;
  uint32_t* ids;
};
void launch_level_pack(const LevelDev& d, const LevelPacked& p, hipStream_t stream);

// Zeroes every manifold slot a pass left unwritten (slot j of pair k with j at least
// min(manifolds[k], cap)), as bm_level_host leaves them.
void launch_level_clear(const LevelDev& d, hipStream_t stream);

}  // namespace bm
