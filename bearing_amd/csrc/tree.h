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

  int insert(int it, const double* b);
  void remove(int leaf);
  void move(int leaf, const double* b);
  template <class F>
  void query(const double* q, F&& visit) {
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

 private:
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
