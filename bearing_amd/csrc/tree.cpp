// src/tree.js, restated on the host for the broadphase (see tree.h).
#include "tree.h"

#include <cmath>

namespace bm {

namespace {
inline double surface(double x0, double y0, double z0, double x1, double y1, double z1) {
  const double x = x1 - x0, y = y1 - y0, z = z1 - z0;
  return x * y + y * z + z * x;
}
// Math.min / Math.max on box coordinates: finite, with -0 below +0 as JavaScript orders them.
inline double mn(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
inline double mx(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}
}  // namespace

// src/tree.js:44-66
int Tree::alloc() {
  if (free == NIL) {
    const int old = (int)parent.size(), capacity = old * 2 > 16 ? old * 2 : 16;
    box.resize(6 * (size_t)capacity);
    parent.resize(capacity);
    child1.resize(capacity);
    child2.resize(capacity);
    height.resize(capacity);
    item.resize(capacity);
    for (int i = old; i < capacity; i++) {
      child1[i] = i + 1 < capacity ? i + 1 : NIL;
      height[i] = -1;
    }
    free = old;
  }
  const int n = free;
  free = child1[n];
  parent[n] = NIL; child1[n] = NIL; child2[n] = NIL;
  height[n] = 0; item[n] = NIL;
  count++;
  return n;
}

void Tree::release(int n) {
  child1[n] = free;
  height[n] = -1;
  free = n;
  count--;
}

int Tree::insert(int it, const double* b) {
  const int leaf = alloc();
  for (int k = 0; k < 6; k++) box[6 * leaf + k] = b[k];
  item[leaf] = it;
  leaves++;
  insert_leaf(leaf);
  return leaf;
}

void Tree::remove(int leaf) {
  remove_leaf(leaf);
  release(leaf);
  leaves--;
}

void Tree::move(int leaf, const double* b) {
  remove_leaf(leaf);
  for (int k = 0; k < 6; k++) box[6 * leaf + k] = b[k];
  insert_leaf(leaf);
}

// src/tree.js:123-165 (Catto's SAH insertion with the branch-and-bound descent cost)
void Tree::insert_leaf(int leaf) {
  if (root == NIL) {
    root = leaf;
    parent[leaf] = NIL;
    return;
  }
  const double l[6] = {box[6 * leaf], box[6 * leaf + 1], box[6 * leaf + 2], box[6 * leaf + 3], box[6 * leaf + 4],
                       box[6 * leaf + 5]};
  int index = root;
  while (height[index] > 0) {
    const double* b = &box[6 * index];
    const double area = surface(b[0], b[1], b[2], b[3], b[4], b[5]);
    const double combined = surface(mn(b[0], l[0]), mn(b[1], l[1]), mn(b[2], l[2]), mx(b[3], l[3]), mx(b[4], l[4]),
                                    mx(b[5], l[5]));
    const double cost = 2 * combined;
    const double inherited = 2 * (combined - area);
    const int c1 = child1[index], c2 = child2[index];
    const double cost1 = descend(c1, l) + inherited;
    const double cost2 = descend(c2, l) + inherited;
    if (cost < cost1 && cost < cost2) break;
    index = cost1 <= cost2 ? c1 : c2;
  }
  const int sibling = index, old_parent = parent[sibling];
  const int p = alloc();
  parent[p] = old_parent;
  fit(p, sibling, leaf);
  height[p] = height[sibling] + 1;
  if (old_parent != NIL) {
    if (child1[old_parent] == sibling) child1[old_parent] = p;
    else child2[old_parent] = p;
  } else {
    root = p;
  }
  child1[p] = sibling; child2[p] = leaf;
  parent[sibling] = p; parent[leaf] = p;
  refit(parent[leaf]);
}

double Tree::descend(int c, const double* l) const {
  const double* b = &box[6 * c];
  const double un = surface(mn(b[0], l[0]), mn(b[1], l[1]), mn(b[2], l[2]), mx(b[3], l[3]), mx(b[4], l[4]), mx(b[5], l[5]));
  if (height[c] == 0) return un;
  return un - surface(b[0], b[1], b[2], b[3], b[4], b[5]);
}

// src/tree.js:176-195
void Tree::remove_leaf(int leaf) {
  if (leaf == root) {
    root = NIL;
    return;
  }
  const int p = parent[leaf], grand = parent[p];
  const int sibling = child1[p] == leaf ? child2[p] : child1[p];
  if (grand != NIL) {
    if (child1[grand] == p) child1[grand] = sibling;
    else child2[grand] = sibling;
    parent[sibling] = grand;
    release(p);
    refit(grand);
  } else {
    root = sibling;
    parent[sibling] = NIL;
    release(p);
  }
}

void Tree::refit(int index) {
  while (index != NIL) {
    index = balance(index);
    const int c1 = child1[index], c2 = child2[index];
    height[index] = 1 + (height[c1] > height[c2] ? height[c1] : height[c2]);
    fit(index, c1, c2);
    index = parent[index];
  }
}

void Tree::fit(int n, int p, int q) {
  double* b = &box[0];
  const int o = 6 * n, i = 6 * p, j = 6 * q;
  b[o] = mn(b[i], b[j]); b[o + 1] = mn(b[i + 1], b[j + 1]); b[o + 2] = mn(b[i + 2], b[j + 2]);
  b[o + 3] = mx(b[i + 3], b[j + 3]); b[o + 4] = mx(b[i + 4], b[j + 4]); b[o + 5] = mx(b[i + 5], b[j + 5]);
}

// src/tree.js:219-252
int Tree::balance(int a) {
  if (height[a] < 2) return a;
  const int b = child1[a], c = child2[a];
  const int bal = height[c] - height[b];
  if (bal > 1) return rotate(a, c, b, false);
  if (bal < -1) return rotate(a, b, c, true);
  return a;
}

int Tree::rotate(int a, int up, int other, bool up_is_first) {
  const int f = child1[up], g = child2[up];
  const int top = parent[a];
  child1[up] = a;
  parent[up] = top;
  parent[a] = up;
  if (top != NIL) {
    if (child1[top] == a) child1[top] = up;
    else child2[top] = up;
  } else {
    root = up;
  }
  const int tall = height[f] > height[g] ? f : g, short_ = tall == f ? g : f;
  child2[up] = tall;
  if (up_is_first) child1[a] = short_;
  else child2[a] = short_;
  parent[short_] = a;
  fit(a, other, short_);
  fit(up, a, tall);
  height[a] = 1 + (height[other] > height[short_] ? height[other] : height[short_]);
  height[up] = 1 + (height[a] > height[tall] ? height[a] : height[tall]);
  return up;
}

}  // namespace bm
