// Internal declarations shared by the host tree build and the HIP solver.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tabi {

// Tree in flat form, nodes in creation (depth-first pre-order) order; the root
// is node 0.  Field meanings follow the reference's tnode
// (src/treecode3d_pb.f:38-47); particle ranges are 0-based inclusive.
struct HostTree {
    int32_t n = 0;
    std::vector<int32_t> ibeg, iend, numpar, level, nchild, child_off;
    std::vector<int32_t> children;                   // flat child lists
    std::vector<double> xmin, xmax, ymin, ymax, zmin, zmax;
    std::vector<double> xmid, ymid, zmid, radius;
    std::vector<int32_t> leaves;                     // leaf node ids, depth-first
    std::vector<int32_t> order;                      // tree position -> original index
    int32_t nnode() const { return (int32_t)ibeg.size(); }
};

// Bit-exact restatement of SETUP + CREATE_TREE + PARTITION_8 + PARTITION
// (src/treecode3d_pb.f:54-356, 1677-1759).  x, y, z are permuted in place.
void build_tree_host(int32_t n, double *x, double *y, double *z, int32_t maxparnode,
                     HostTree &t);

}  // namespace tabi
