// hmatrix_host.hpp -- cluster tree and block partition of the H-matrix
// operator (see hmatrix_host.cpp).
#pragma once
#include <vector>

namespace hamsx {

struct HNode {
  int s, e;          // cluster range [s, e] in the permuted (cluster) order
  int left, right;   // children, -1 for a leaf
  double c[3];       // centroid of the cluster's panel centres
  double diam;       // bounding-box diagonal
};

// One block of the partition: rows [rs, rs + nr) x columns [cs, cs + nc) in
// cluster order; aca = 1 for an admissible block that is compressed by ACA
// (it may still fall back to dense when its rank does not pay off), 0 for a
// block stored dense.
struct HBlk {
  int rs, nr, cs, nc, aca;
};

// perm[k] = panel index (0-based) at cluster position k; blocks in the
// reference's creation order.  cen: n x 3, column-major.
void hmatrix_partition(int n, const double* cen, int leaf, double eta, std::vector<int>& perm,
                       std::vector<HNode>& tree, std::vector<HBlk>& blocks);

}  // namespace hamsx
