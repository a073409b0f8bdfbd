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

// The same tree built on the GPU (level by level, data-parallel partition).
// d_x, d_y, d_z: centroids in original order on the device; d_tx, d_ty, d_tz
// receive them in tree order and d_order the 0-based permutation.  t receives
// the node table and order exactly as build_tree_host produces them.
void build_tree_gpu(int32_t n, const double *d_x, const double *d_y, const double *d_z,
                    int32_t maxparnode, void *stream, HostTree &t, double *d_tx,
                    double *d_ty, double *d_tz, int32_t *d_order);

// Device memory from the library's caching allocator.
void *device_alloc(size_t bytes);
void device_free(void *p);

}  // namespace tabi
