// The broadphase's dynamic AABB tree, on the host: the order the pair list fills in
// depends on the tree's exact shape (src/tree.js: SAH insertion 123-165, AVL rotations
// 219-252, child1-first depth-first query 259-275), so it is kept exactly as the
// reference grows, balances and walks it. Node storage, the free list and the order
// nodes are handed out follow src/tree.js:44-76 too.
#pragma once
#include <cstdint>
#include <vector>

namespace bm {

class Tree {
 public:
  static constexpr int NIL = -1;
  int root = NIL, count = 0, leaves = 0, free = NIL;
  std::vector<double> box;  // 6 per node: min x y z, max x y z
  std::vector<int> parent, child1, child2, height, item;
  std::vector<int> stack;

  std::vector<double> at;

  int insert(int it, const double* b);
  void remove(int leaf);
  void move(int leaf, const double* b);
  // src/tree.js:259-275 (query): the leaves whose boxes overlap q, child1's side first. The
  // walk's stack is the tree's own, or the caller's (so that several threads can walk one
  // tree at once).
  template <class F>
  void query(const double* q, F&& visit) {
    query(q, visit, stack);
  }
  template <class F>
  void query(const double* q, F&& visit, std::vector<int>& stack) const {
    if (root == NIL) return;
    stack.clear();
    stack.push_back(root);
    while (!stack.empty()) {
      const int n = stack.back();
      stack.pop_back();
      const double* b = &box[6 * n];
      if (b[0] > q[3] || b[3] < q[0] || b[1] > q[4] || b[4] < q[1] || b[2] > q[5] || b[5] < q[2]) continue;
      if (height[n] == 0) {
        visit(item[n]);
        continue;
      }
      stack.push_back(child2[n]);
      stack.push_back(child1[n]);
    }
  }

  // src/tree.js:288-334 (raycast): the ray from o along d, boxes grown by g, nearer child
  // first; visit(item) returns the new maxT. Returns the last maxT. The walk's stack is the
  // tree's own, or the caller's (so that several threads can walk one tree at once).
  template <class F>
  double raycast(const double* o, const double* d, double max_t, const double* g, F&& visit) {
    return raycast(o, d, max_t, g, visit, stack, at);
  }
  template <class F>
  double raycast(const double* o, const double* d, double max_t, const double* g, F&& visit, std::vector<int>& stack,
                 std::vector<double>& at) const {
    if (root == NIL) return max_t;
    const double ix = 1 / d[0], iy = 1 / d[1], iz = 1 / d[2];
    const double inv[3] = {ix, iy, iz};
    const double t = enter(root, o, d, inv, g, max_t);
    if (t < 0) return max_t;
    stack.clear();
    at.clear();
    stack.push_back(root);
    at.push_back(t);
    while (!stack.empty()) {
      const int n = stack.back();
      const double tn = at.back();
      stack.pop_back();
      at.pop_back();
      if (tn > max_t) continue;
      if (height[n] == 0) {
        max_t = visit(item[n]);
        continue;
      }
      const int c1 = child1[n], c2 = child2[n];
      const double t1 = enter(c1, o, d, inv, g, max_t), t2 = enter(c2, o, d, inv, g, max_t);
      if (t1 >= 0 && t2 >= 0) {
        if (t2 < t1) {
          stack.push_back(c1); at.push_back(t1); stack.push_back(c2); at.push_back(t2);
        } else {
          stack.push_back(c2); at.push_back(t2); stack.push_back(c1); at.push_back(t1);
        }
      } else if (t1 >= 0) {
        stack.push_back(c1); at.push_back(t1);
      } else if (t2 >= 0) {
        stack.push_back(c2); at.push_back(t2);
      }
    }
    return max_t;
  }

 private:
  // src/tree.js:336-367 (enter): where the ray enters node n's box grown by g, or -1.
  double enter(int n, const double* o, const double* d, const double* inv, const double* g, double max_t) const {
    const double* b = &box[6 * n];
    double t0 = 0, t1 = max_t;
    for (int k = 0; k < 3; k++) {
      if (d[k] == 0) {
        if (o[k] < b[k] - g[k] || o[k] > b[k + 3] + g[k]) return -1;
      } else {
        double near = (b[k] - g[k] - o[k]) * inv[k], far = (b[k + 3] + g[k] - o[k]) * inv[k];
        if (near > far) { const double s = near; near = far; far = s; }
        if (near > t0) t0 = near;
        if (far < t1) t1 = far;
        if (t0 > t1) return -1;
      }
    }
    return t0;
  }
  int alloc();
  void release(int n);
  void insert_leaf(int leaf);
  void remove_leaf(int leaf);
  double descend(int c, const double* l) const;
  void refit(int index);
  void fit(int n, int p, int q);
  int balance(int a);
  int rotate(int a, int up, int other, bool up_is_first);
};

}  // namespace bm
