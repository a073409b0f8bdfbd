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

// Learned initial guess (warmstart.hip).  Packed parameter layout (floats):
// feature means [12], feature scales [12], output scales [2], 6 unused, then
// W1 [128][12], b1 [128], W2 [128][128], b2, W3 [128][128], b3, W4 [2][128], b4 [2].
namespace ws {
constexpr int kF = 12, kH = 128, kO = 2;
constexpr int oMu = 0, oSig = 12, oTs = 24, oW1 = 32;
constexpr int oB1 = oW1 + kH * kF, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kO * kH, kNP = oB4 + kO;
}  // namespace ws

struct WsFeatIn {
    int n = 0;                                        // faces (tree order)
    const double *tx, *ty, *tz, *tnx, *tny, *tnz, *tarea;
    const int32_t *leaf_of, *leaf_beg, *leaf_m2;      // leaf of each face; first face
                                                      // and 2*size of each leaf
    int nq = 0;
    const double *qpos;                               // [3nq]
    const double *q;                                  // [nq]
    double eps_p = 1.0;
    const double *b;                                  // right-hand side [2n]
    double scale = 1.0;                               // 1/rms(b[0..n-1])
};
struct WsModel;
WsModel *ws_create(const float *params);
void ws_destroy(WsModel *m);
void ws_get_params(WsModel *m, float *out);
void ws_features(const WsFeatIn &in, float *d_feat, void *stream);
void ws_predict_dev(WsModel *m, int64_t n, const float *d_feat, double scale, double *d_x0,
                    void *stream);
void ws_predict(WsModel *m, int64_t n, const float *feat, double scale, double *x0);
void ws_grad(WsModel *m, int64_t ns, const float *feat, const float *target, float *grad_out,
             double *loss_out);
void ws_train(WsModel *m, int64_t ns, const float *feat, const float *target, int epochs,
              int batch, float lr, uint64_t seed, float *epoch_loss);

// Device memory from the library's caching allocator.
void *device_alloc(size_t bytes);
void device_free(void *p);

}  // namespace tabi
