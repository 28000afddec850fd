// ORACLE — test infrastructure only (see soft.cpp).
//
// src/graph.js:105-170 (colour) and the row bookkeeping of arrange (178-188),
// line for line: greedy lowest-free-bit colouring in manifold order over 32-bit
// masks, the overflow and resting runs, stable counting sorts by b's slot (two
// 12-bit digits) and then by colour, the run starts, and the way back (row).
// Pinned by the colour() calls recorded from the reference (tests/golden, colour.*).
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace {
constexpr int DYNAMIC = 0, STATIC = 1;
constexpr int COLOURS = 32, OVERFLOW = COLOURS, RESTING = OVERFLOW + 1;
constexpr int DIGIT = 12, DIGITS = 1 << DIGIT;

int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
}  // namespace

extern "C" int oracle_colour(int32_t count, const int32_t* A, const int32_t* B, int32_t top, const uint8_t* type,
                             const uint8_t* frozen, int32_t frozen_count, int32_t* colour, int32_t* order,
                             int32_t* row, int32_t* starts, int32_t* solved) {
  std::vector<int32_t> used(top > 0 ? top : 1, 0);
  const int n = count;
  const bool any_frozen = frozen_count > 0;
  for (int m = 0; m < n; m++) { used[A[m]] = 0; used[B[m]] = 0; }
  for (int m = 0; m < n; m++) {
    const int a = A[m], b = B[m];
    bool da = type[a] == DYNAMIC, db = type[b] == DYNAMIC;
    if (any_frozen) {
      const bool fa = frozen[a] != 0, fb = frozen[b] != 0;
      if ((fa || type[a] == STATIC) && (fb || type[b] == STATIC)) { colour[m] = RESTING; continue; }
      if (fa) da = false;
      if (fb) db = false;
    }
    const int32_t held = (da ? used[a] : 0) | (db ? used[b] : 0);
    if (held == -1) { colour[m] = OVERFLOW; continue; }
    const int32_t bit = ~held & (held + 1);  // int32 wrap as JavaScript's |0
    colour[m] = 31 - clz32((uint32_t)bit);
    if (da) used[a] |= bit;
    if (db) used[b] |= bit;
  }
  // Stable counting sorts: by b's low digit, its high digit past 4,096 slots, then colour.
  std::vector<int32_t> cnt(DIGITS + 1);
  const bool high = top > DIGITS;
  int32_t* from = high ? order : row;
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int m = 0; m < n; m++) cnt[(B[m] & (DIGITS - 1)) + 1]++;
  for (int k = 0; k < DIGITS; k++) cnt[k + 1] += cnt[k];
  for (int m = 0; m < n; m++) from[cnt[B[m] & (DIGITS - 1)]++] = m;
  if (high) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = 0; i < n; i++) cnt[((uint32_t)B[order[i]] >> DIGIT) + 1]++;
    for (int k = 0; k < DIGITS; k++) cnt[k + 1] += cnt[k];
    for (int i = 0; i < n; i++) { const int m = order[i]; row[cnt[(uint32_t)B[m] >> DIGIT]++] = m; }
    from = row;
  }
  for (int k = 0; k < COLOURS + 3; k++) starts[k] = 0;
  for (int m = 0; m < n; m++) starts[colour[m] + 1]++;
  for (int k = 0; k <= RESTING; k++) starts[k + 1] += starts[k];
  for (int i = 0; i < n; i++) { const int m = from[i]; order[starts[colour[m]]++] = m; }
  for (int k = RESTING; k > 0; k--) starts[k] = starts[k - 1];
  starts[0] = 0;
  *solved = starts[RESTING];
  for (int o = 0; o < n; o++) row[order[o]] = o;  // arrange's way back, src/graph.js:187
  return 0;
}
