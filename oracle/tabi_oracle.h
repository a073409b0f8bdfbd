/* CPU oracle for the TABI-PB treecode-accelerated GMRES solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (gengwh/TABI-PB, src/treecode3d_pb.f, src/main.f90, src/readin.f90,
 * src/dgmres_dep.f) used by tests/ and by bench.py's CPU-baseline leg as the
 * checker.  No product code path links, loads or calls it.
 *
 * Conventions: every array is 0-based; "tree order" is the order after the
 * tree build's in-place permutation (src/main.f90:447-453); vectors of length
 * 2n are [phi(0..n-1); dphi/dn(0..n-1)].  Points are interleaved xyz (n x 3,
 * the memory layout of the driver's tr_xyz(3,n)).
 */
#ifndef TABI_ORACLE_H
#define TABI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* node record of the restated tree (one per tnode, src/treecode3d_pb.f:38-47) */
typedef struct {
    int32_t ibeg, iend, numpar, level, nchild;
    int32_t child[8];
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double xmid, ymid, zmid, radius;
} or_node;

typedef struct {
    double eps;      /* eps_w / eps_p */
    double kappa;    /* inverse Debye length */
    double theta;    /* MAC parameter */
    int32_t order;   /* Taylor order p */
    int32_t maxparnode;
} or_params;

/* kappa exactly as the driver forms it, src/main.f90:62-63 */
double or_kappa(double bulk_strength, double eps_w);

/* centroids / normals / Heron areas of faces (form_matrix, src/main.f90:569-594);
 * faces are 0-based vertex triples */
void or_face_geometry(int32_t nface, const double *vpos, const double *vnrm,
                      const int32_t *faces, double *xyz, double *nrm, double *area);

/* surface_area face culling (src/readin.f90:264-333): keeps faces whose Heron
 * area is >= 1e-5 and whose centroid is farther than squared distance 1e-5 from
 * each of the 10 preceding faces (original numbering).  Writes kept faces to
 * faces_out and returns their count; *s_area gets the summed area of all faces. */
int32_t or_cull_faces(int32_t nface, const double *vpos, const int32_t *faces,
                      int32_t *faces_out, double *s_area);

/* right-hand side (src/main.f90:597-607), in the order of xyz/nrm */
void or_rhs(int32_t n, const double *xyz, const double *nrm, int32_t nq,
            const double *qpos, const double *q, double eps_p, double *b);

/* tree build (SETUP + CREATE_TREE + PARTITION_8 + PARTITION,
 * src/treecode3d_pb.f:54-356,1677-1759).  xyz is permuted in place into tree
 * order, orderarr[i] receives the 0-based original index of tree position i.
 * nodes must hold 2n+1 records; returns the node count (nodes[0] is the root,
 * nodes are numbered in creation = depth-first pre-order). */
int32_t or_build_tree(int32_t n, double *xyz, int32_t maxparnode,
                      int32_t *orderarr, or_node *nodes);

/* leaves in depth-first order; returns count; ibeg/iend 0-based inclusive */
int32_t or_leaves(const or_node *nodes, int32_t *ibeg, int32_t *iend);

/* treecode matvec y = A x (TREE_COMPP_PB, src/treecode3d_pb.f:1130-1189),
 * inputs in tree order.  counts (may be NULL) receives
 * {near pairs excluding self, accepted far pairs, sum of numpar over nodes whose
 * moments were formed}. */
void or_treecode_matvec(const or_params *prm, int32_t n, const double *xyz,
                        const double *nrm, const double *area, const or_node *nodes,
                        const double *x, double *y, int64_t *counts);

/* the same for a subset of targets (tree positions targets[0..nt-1]);
 * y receives {(Ax)_i, (Ax)_{n+i}} interleaved per listed target */
void or_treecode_matvec_targets(const or_params *prm, int32_t n, const double *xyz,
                                const double *nrm, const double *area, const or_node *nodes,
                                const double *x, int32_t nt, const int32_t *targets,
                                double *y, int64_t *counts);

/* O(n^2) direct-sum matvec with the same kernel and self exclusion */
void or_direct_matvec(const or_params *prm, int32_t n, const double *xyz,
                      const double *nrm, const double *area, const double *x, double *y);

/* block-diagonal leaf preconditioner z = M^-1 r (LEAFMATVECpara,
 * src/treecode3d_pb.f:1393-1542), fresh partial-pivot LU per leaf.
 * Returns 0, or 1+leaf index of the first singular block. */
int32_t or_precond(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                   const double *area, int32_t nleaf, const int32_t *ibeg,
                   const int32_t *iend, const double *r, double *z);

/* dense leaf block (column-major 2m x 2m) exactly as LEAFMATVECpara lays it out */
void or_leaf_block(const or_params *prm, const double *xyz, const double *nrm,
                   const double *area, int32_t ibeg, int32_t iend, double *blk);

/* restarted GMRES with left preconditioning and ITOL=2 (DGMRES + DPIGMR +
 * DORTH + DHEQR + DHELS + DRLCAL + ISDGMR, src/dgmres_dep.f), using the treecode
 * matvec and the leaf preconditioner above.  x holds the initial guess on entry.
 * Returns ierr (0 converged, 1 iteration cap, 2 stalled); iter/err like DGMRES.
 * err_hist (may be NULL, length >= (nrmax+1)*maxl) gets each printed err. */
int32_t or_gmres(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                 const double *area, const or_node *nodes, const double *b, double *x,
                 double tol, int32_t maxl, int32_t nrmax, int32_t *iter, double *err,
                 double *err_hist);

/* solvation energy (kcal/mol) from a tree-order solution
 * (src/main.f90:144-157 with potential_molecule src/main.f90:614-657);
 * couleng receives the Coulomb energy term. */
double or_energy(int32_t n, const double *xyz, const double *nrm, const double *area,
                 const double *x, int32_t nq, const double *qpos, const double *q,
                 double eps, double kappa, double eps_p, double *couleng);

/* Taylor coefficients a_k of exp(-kappa|d|)/|d| for |k| <= lim (COMP_TCOEFF,
 * src/treecode3d_pb.f:684-855); out is (lim+1)^3 with index (i*(lim+1)+j)*(lim+1)+k */
void or_taylor_coeffs(double dx, double dy, double dz, double kappa, int32_t lim,
                      double *out);

/* rows of the O(n^2) direct operator for the listed targets (y[2k], y[2k+1]) */
void or_direct_rows(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                    const double *area, const double *x, int32_t nt, const int32_t *targets,
                    double *y);
/* MSOLVEprec (src/main.f90:532-542) */
void or_diag_prec(int32_t n2, double eps, const double *r, double *z);
/* output_potential vertex values (src/main.f90:226-326), original face order */
int32_t or_vertex_potential(int32_t n, int32_t nvert, const double *vpos, const int32_t *faces1,
                            const double *xyz, const double *x, double *vpot);

#ifdef __cplusplus
}
#endif
#endif
