// hmatrix_host.cpp -- the cluster tree and block partition of the H-matrix
// operator (host side; the ACA compression and the H-matvec run on the
// device, hmatrix.hip).
//
// Geometric bisection of the panel centroids and the eta-admissibility
// partition of the reference, HMatrix.f90:217-533 (SPLIT_CLUSTER_ITERATIVE,
// BUILD_BLOCK_TREE_ITERATIVE): the same explicit-stack traversals, split
// planes, partition sweep and admissibility test, so the permutation and
// the block list come out in the reference's order.
#include <cmath>
#include <vector>

#include "hmatrix_host.hpp"

namespace hamsx {

void hmatrix_partition(int n, const double* cen, int leaf, double eta, std::vector<int>& perm,
                       std::vector<HNode>& tree, std::vector<HBlk>& blocks) {
  perm.resize(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  tree.clear();
  blocks.clear();
  if (n <= 0) return;
  // working copy of the coordinates, permuted along with perm
  std::vector<double> w((size_t)3 * n);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) w[3 * (size_t)i + d] = cen[i + (size_t)n * d];

  // ---- cluster tree (:217-371) -------------------------------------------
  struct Item {
    int is, ie, parent, side;
  };
  std::vector<Item> st;
  st.push_back({0, n - 1, -1, 0});
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    const int ns = it.ie - it.is + 1;
    if (ns <= 0) continue;
    const int id = (int)tree.size();
    tree.push_back(HNode{});
    if (it.parent >= 0) (it.side == 1 ? tree[it.parent].left : tree[it.parent].right) = id;
    HNode& nd = tree[id];
    nd.s = it.is;
    nd.e = it.ie;
    nd.left = nd.right = -1;
    double bmin[3], bmax[3], sum[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; d++) bmin[d] = bmax[d] = w[3 * (size_t)it.is + d];
    for (int j = it.is; j <= it.ie; j++)
      for (int d = 0; d < 3; d++) {
        const double v = w[3 * (size_t)j + d];
        if (v < bmin[d]) bmin[d] = v;
        if (v > bmax[d]) bmax[d] = v;
        sum[d] += v;
      }
    double len[3];
    for (int d = 0; d < 3; d++) {
      nd.c[d] = sum[d] / (double)ns;
      len[d] = bmax[d] - bmin[d];
    }
    nd.diam = std::sqrt(len[0] * len[0] + len[1] * len[1] + len[2] * len[2]);
    if (ns <= leaf) continue;
    int sd = 0;
    double ml = len[0];
    if (len[1] > ml) {
      sd = 1;
      ml = len[1];
    }
    if (len[2] > ml) sd = 2;
    const double split = 0.5 * (bmin[sd] + bmax[sd]);
    int i = it.is, j = it.ie;
    while (i <= j) {
      if (w[3 * (size_t)i + sd] <= split) {
        i++;
      } else {
        for (int d = 0; d < 3; d++) std::swap(w[3 * (size_t)i + d], w[3 * (size_t)j + d]);
        std::swap(perm[i], perm[j]);
        j--;
      }
    }
    int le = i - 1;
    if (le < it.is) le = it.is;
    if (le >= it.ie) le = it.ie - 1;
    st.push_back({le + 1, it.ie, id, 2});
    st.push_back({it.is, le, id, 1});
  }

  // ---- block partition (:379-533) ----------------------------------------
  std::vector<std::pair<int, int>> bs;
  bs.push_back({0, 0});
  while (!bs.empty()) {
    const int r = bs.back().first, c = bs.back().second;
    bs.pop_back();
    const HNode &R = tree[r], &C = tree[c];
    const int nr = R.e - R.s + 1, nc = C.e - C.s + 1;
    if (nr <= 0 || nc <= 0) continue;
    const bool lr = R.left < 0, lc = C.left < 0;
    const double dx = R.c[0] - C.c[0], dy = R.c[1] - C.c[1], dz = R.c[2] - C.c[2];
    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double dmax = std::max(R.diam, C.diam);
    if (dist > eta * dmax && nr > 1 && nc > 1) {
      // ADD_LOWRANK_BLOCK keeps blocks with a side of 4 or fewer dense (:624-627)
      blocks.push_back(HBlk{R.s, nr, C.s, nc, (nr > 4 && nc > 4) ? 1 : 0});
      continue;
    }
    if (lr && lc) {
      blocks.push_back(HBlk{R.s, nr, C.s, nc, 0});
      continue;
    }
    if (lr) {
      bs.push_back({r, C.right});
      bs.push_back({r, C.left});
    } else if (lc) {
      bs.push_back({R.right, c});
      bs.push_back({R.left, c});
    } else {
      bs.push_back({R.right, C.right});
      bs.push_back({R.right, C.left});
      bs.push_back({R.left, C.right});
      bs.push_back({R.left, C.left});
    }
  }
}

}  // namespace hamsx
