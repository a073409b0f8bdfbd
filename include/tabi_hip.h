/* tabi_hip.h -- C ABI of the MI355X (gfx950) treecode-accelerated boundary-integral
 * Poisson-Boltzmann solver (libtabi_hip.so).
 *
 * The library replaces the treecode matvec, the leaf-block preconditioner and
 * the set-up work of gengwh/TABI-PB behind the two callbacks its SLATEC DGMRES
 * calls.  All vectors are contiguous float64 in TREE ORDER, laid out as
 * [phi(1..n); dphi/dn(1..n)] with n = number of faces (the reference's
 * N = 2*nface, src/main.f90:84).  Points are interleaved x,y,z per face (the
 * memory layout of the Fortran arrays tr_xyz(3,n) / tr_q(3,n)).
 *
 * Every function returns 0 on success and a nonzero status on failure; the
 * message is available from tabi_last_error().  The library holds one default
 * solver context (the drop-in path) plus any number of explicit handles.
 * Calls on one context are not re-entrant (the reference's matvec mutates
 * global state too, src/treecode3d_pb.f:1169-1182).
 */
#ifndef TABI_HIP_H
#define TABI_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Solver parameters.  eps = eps_w/eps_p (src/main.f90:65); kappa as formed at
 * src/main.f90:62-63; order = Taylor order p, maxparnode = leaf size, theta = MAC
 * (usrdata.in lines 6-8, src/main.f90:45-47). */
typedef struct {
    double eps;
    double kappa;
    double theta;
    int32_t order;       /* 1..6 */
    int32_t maxparnode;  /* >= 1 */
    int32_t precision;   /* 0: FP64 throughout; 1: mixed (FP32 near/far field with
                            compensated per-target sums, FP64 elsewhere) */
    int32_t flags;       /* bit 0: preconditioner with matvec-consistent L2/L3
                            placement instead of the reference's (tc.f:1481-1488);
                            bit 1: build the tree with the sequential host
                            partition instead of the GPU one (same result);
                            bit 2: diagonal-scaling preconditioner MSOLVEprec
                            (src/main.f90:532-542, usrdata.in iprcd != 1)
                            instead of the leaf-block LU (no factorisation);
                            bit 4: GMRES with its control on the device, one
                            captured graph per restart cycle (same results);
                            bit 5 (mixed precision): round the finished FP64
                            leaf factors to FP32 in one pass instead of as each
                            entry is finished (same values);
                            bits 8-11 (t): field kernels cut the work into about
                            2^(10+t) tasks per launch (t = 0: 2^16) */
} tabi_params;

typedef struct tabi_ctx tabi_ctx;

/* ---- default-context drop-in API (what the Fortran driver binds) ---------- */

/* Set-up for one surface (replaces treecode_initialization, src/main.f90:329-464,
 * and SETUP/CREATE_TREE, src/treecode3d_pb.f:54-356).  Inputs in ORIGINAL face
 * order: centroids xyz[3n], unit normals nrm[3n], areas area[n] (as form_matrix
 * computes them, src/main.f90:569-594).  Builds the tree, the per-target
 * interaction lists and the factored leaf blocks on the GPU.  orderarr_out[n]
 * receives the 1-based tree permutation (tree position i holds original face
 * orderarr(i), src/treecode3d_pb.f:30); nleaf_out the leaf count. */
int tabi_init(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
              const double *area, int32_t *orderarr_out, int32_t *nleaf_out);

/* y = A x, the MATVEC callback (MATVECmain, src/main.f90:467-488, which calls
 * TREE_COMPP_PB, src/treecode3d_pb.f:1130-1189).  Host pointers, n2 = 2n.
 * Reproduces the shortcut y = x when sum|x| < 1e-10 (src/main.f90:478-479). */
int tabi_matvec(int32_t n2, const double *x, double *y);

/* z = M^-1 r, the MSOLVE callback (MSOLVEdric, src/main.f90:511-529, which calls
 * LEAFMATVECpara, src/treecode3d_pb.f:1393-1542).  Host pointers. */
int tabi_msolve(int32_t n2, const double *r, double *z);

/* Right-hand side in tree order (form_matrix, src/main.f90:597-607) for nq point
 * charges q at qpos[3nq], interior dielectric eps_p. */
int tabi_rhs(int32_t nq, const double *qpos, const double *q, double eps_p, double *b);

/* Solvation and Coulomb energies in kcal/mol from a tree-order solution x
 * (src/main.f90:144-157, potential_molecule src/main.f90:614-657). */
int tabi_energy(int32_t nq, const double *qpos, const double *q, double eps_p,
                const double *x, double *soleng, double *couleng);

/* Vertex-averaged surface potential (output_potential, src/main.f90:226-326):
 * for each of nvert vertices vpos[3*nvert], the mean of the tree-order solution
 * x[2n] over the first 14 faces in the original numbering that use the vertex,
 * weighted by 1/|centroid - vertex|, scaled by para*4*pi (src/main.f90:245,
 * 289): vpot[2*nvert] = (phi, dphi/dn) per vertex, as surface_potential.dat
 * holds them; a vertex no face uses gets NaN, as in the reference.  faces[3n]
 * are the surface's (culled) faces in the original order, 1-based vertex
 * indices (the nvert array of the reference). */
int tabi_vertex_potential(int32_t nvert, const double *vpos, const int32_t *faces,
                          const double *x, double *vpot);

/* Writes surface_potential.dat in the reference's layout (src/main.f90:307-316):
 * "nspt nface", per vertex (i10,6f12.6,2f20.10) of index, position, normal and
 * vpot, then each face's three 1-based vertex indices. */
int tabi_write_surface_potential(const char *path, int32_t nvert, const double *vpos,
                                 const double *vnrm, const double *vpot, int32_t nface,
                                 const int32_t *faces);

/* GPU-resident restarted GMRES with the reference's settings and stopping rule
 * (DGMRES with ITOL=2, JPRE=-1, JSCAL=0, src/dgmres_dep.f:279-852; MAXL/NRMAX
 * from src/main.f90:93-97).  b and x (initial guess in, solution out) are host
 * pointers in tree order; iter/err/ierr as DGMRES returns them. */
int tabi_gmres(const double *b, double *x, double tol, int32_t maxl, int32_t nrmax,
               int32_t *iter, double *err, int32_t *ierr);

/* Releases the default context. */
void tabi_finalize(void);

/* Last error message of the calling thread ("" if none). */
const char *tabi_last_error(void);

/* Counters of the default context: {nface, nnode, nleaf, P_near, P_far, P_mom,
 * n_far_items, n_near_items, n_tasks, lu_bytes}. */
int tabi_stats(int64_t *out10);

/* ---- explicit-handle API (batched and device-resident use) ---------------- */

int tabi_ctx_create(const tabi_params *prm, int32_t n, const double *xyz,
                    const double *nrm, const double *area, int32_t *orderarr_out,
                    tabi_ctx **out);
void tabi_ctx_destroy(tabi_ctx *c);
/* One context over nsurf surfaces ("packed"; no counterpart in the reference,
 * which holds one surface per run): surface k's faces occupy [face0[k],
 * face0[k] + nface[k]) of a common tree-order numbering (face0[k] a multiple of
 * 64; face0_out receives nsurf + 1 entries, the last = n, the packed length),
 * the gaps being padding faces that are never read.  Each surface keeps its own
 * tree and lists (no interaction crosses surfaces); tabi_ctx_matvec /
 * tabi_ctx_msolve then act on packed vectors [first halves of all surfaces (n);
 * second halves (n)] with one launch per kernel for every surface, and give
 * each surface's entries exactly as its own context does. */
int tabi_ctx_create_packed(const tabi_params *prm, int32_t nsurf, const int32_t *nface,
                           const double *const *xyz, const double *const *nrm,
                           const double *const *area, int32_t *face0_out, tabi_ctx **out);
int tabi_ctx_stats(tabi_ctx *c, int64_t *out10);

/* device-pointer matvec / preconditioner, enqueued on the context's own stream
 * and returning before completion; the caller must have finished writing the
 * input on its streams, and must call tabi_ctx_sync before reading the output */
int tabi_ctx_matvec_dev(tabi_ctx *c, const double *d_x, double *d_y);
int tabi_ctx_msolve_dev(tabi_ctx *c, const double *d_r, double *d_z);
int tabi_ctx_sync(tabi_ctx *c);
/* host-pointer forms */
int tabi_ctx_matvec(tabi_ctx *c, const double *x, double *y);
int tabi_ctx_msolve(tabi_ctx *c, const double *r, double *z);
int tabi_ctx_rhs(tabi_ctx *c, int32_t nq, const double *qpos, const double *q,
                 double eps_p, double *b);
int tabi_ctx_energy(tabi_ctx *c, int32_t nq, const double *qpos, const double *q,
                    double eps_p, const double *x, double *soleng, double *couleng);
int tabi_ctx_gmres(tabi_ctx *c, const double *b, double *x, double tol, int32_t maxl,
                   int32_t nrmax, int32_t *iter, double *err, int32_t *ierr);
int tabi_ctx_vertex_potential(tabi_ctx *c, int32_t nvert, const double *vpos,
                              const int32_t *faces, const double *x, double *vpot);

/* tabi_ctx_energy without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * charges d_qpos[3*nq], d_q[nq] and the tree-order solution d_x[2n] are device
 * pointers; the per-call scratch (energy partials, per-charge potentials,
 * Coulomb partials: about 8*(nq/64*n/32 + nq + (nq/256)^2) bytes) lives in a
 * static device array of 32 MiB held by every GPU the library runs on; larger
 * problems are rejected with an error (use tabi_ctx_energy).  Calls that
 * overlap on one GPU share that array and race.  Returns after completion. */
int tabi_ctx_energy_noalloc(tabi_ctx *c, int32_t nq, const double *d_qpos, const double *d_q,
                            double eps_p, const double *d_x, double *soleng, double *couleng);

/* ---- blocks of right-hand sides on one surface ------------------------------
 * Several vectors (charge sets, Krylov vectors) share one tree pass: the near-
 * field kernel values and far-field Taylor expansions are formed once per pair
 * and applied to every vector, and the leaf factors are read once per group of
 * vectors.  Per vector the arithmetic is that of the single-vector entry points
 * (column k of a block result equals the single result for vector k).  Vectors
 * are stored one after another: x[k*2n .. k*2n+2n-1] is vector k, tree order.
 * No counterpart in the reference, which solves one charge set per run. */
int tabi_ctx_matvec_block(tabi_ctx *c, int32_t nrhs, const double *x, double *y);
int tabi_ctx_msolve_block(tabi_ctx *c, int32_t nrhs, const double *r, double *z);
/* nrhs runs of tabi_ctx_gmres (initial guesses in x) in lockstep; iter, err and
 * ierr are per run */
int tabi_ctx_gmres_block(tabi_ctx *c, int32_t nrhs, const double *b, double *x, double tol,
                         int32_t maxl, int32_t nrmax, int32_t *iter, double *err, int32_t *ierr);
/* Solves for nrhs charge sets at common positions qpos[3nq] with charges
 * q[k*nq .. k*nq+nq-1] of set k: right-hand sides, block GMRES from zero
 * (MAXL = NRMAX = 10, tol), energies; x_out[nrhs*2n] may be NULL. */
int tabi_ctx_solve_block(tabi_ctx *c, int32_t nrhs, int32_t nq, const double *qpos,
                         const double *q, double eps_p, double tol, double *x_out, int32_t *iter,
                         double *err, int32_t *ierr, double *soleng, double *couleng);

/* Full solves of nsurf independent surfaces (face arrays xyz[k], nrm[k],
 * area[k] of nface[k] faces, original order; charges qpos[k], q[k] of nq[k]
 * charges), all with the parameters prm: the surfaces are set up from several
 * host threads, then their GMRES runs (MAXL = NRMAX = 10, tol, zero start)
 * advance in lockstep; each step's matvec and preconditioner solve of every
 * active surface is one replay of a captured HIP graph with one branch per
 * surface.  options bit 0 launches the same work eagerly instead of through the
 * graph.  Per-surface results: iter, err, ierr, soleng, couleng.  No
 * counterpart in the reference, which solves one surface per run. */
int tabi_solve_many(const tabi_params *prm, int32_t nsurf, const int32_t *nface,
                    const double *const *xyz, const double *const *nrm,
                    const double *const *area, const int32_t *nq, const double *const *qpos,
                    const double *const *q, double eps_p, double tol, int32_t options,
                    int32_t *iter, double *err, int32_t *ierr, double *soleng, double *couleng);
/* Of the last tabi_solve_many on this thread: {graph captures, graph launches,
 * set-up + right-hand sides ms, GMRES ms, energies ms, total ms}. */
int tabi_last_solve_many_stats(double *out6);
/* Of the last packed tabi_solve_many on this thread, the set-up's parts in ms:
 * {trees and geometry, interaction lists and tasks, leaf LU, right-hand sides}. */
int tabi_last_solve_many_setup(double *out4);

/* The library keeps freed device blocks for reuse by later contexts and solves;
 * this returns all cached (unused) blocks to the runtime. */
int tabi_release_memory(void);

/* Device-memory plumbing for callers that hold no other GPU runtime (so the
 * device-pointer entry points can be used from C, Fortran or ctypes). */
int tabi_device_alloc(int64_t bytes, void **out);
int tabi_device_free(void *p);
int tabi_copy_to_device(void *dst, const void *src, int64_t bytes);
int tabi_copy_to_host(void *dst, const void *src, int64_t bytes);

/* Per-kernel device timing on the context's stream.  With profiling enabled the
 * library brackets each launch of the timed kernels with HIP events; times
 * accumulate until reset.  Slots: 0 moments, 1 near field, 2 far field,
 * 3 reduction, 4 preconditioner solve, 5 list build, 6 LU factorisation,
 * 7 right-hand side + energy.  out_ms[8] gets the summed milliseconds,
 * out_launches[8] the launch counts. */
int tabi_ctx_profile(tabi_ctx *c, int32_t enable);
int tabi_ctx_kernel_times(tabi_ctx *c, double *out_ms, int64_t *out_launches);

/* One full PB solve for a surface already in HBM-resident form: builds the tree,
 * interaction lists and leaf factors, forms the right-hand side, runs GMRES from
 * a zero start to tol, and evaluates the energies.  Host inputs in original face
 * order; x_out[2n] (may be NULL) receives the tree-order solution, orderarr_out
 * (may be NULL) the 1-based permutation. */
int tabi_solve(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
               const double *area, int32_t nq, const double *qpos, const double *q,
               double eps_p, double tol, double *x_out, int32_t *orderarr_out,
               int32_t *iter, double *err, int32_t *ierr, double *soleng,
               double *couleng);

/* Per-phase device times (ms) of the last tabi_solve / tabi_solve_ws on this
 * thread: {tree, lists, precond_factor, rhs, gmres, energy, total, warmstart}. */
int tabi_last_solve_times(double *out8);

/* ---- learned initial guess (no counterpart in the reference) -------------
 * A per-face MLP, 12 features -> 3 x 128 GELU -> 2 outputs, BF16 weights and
 * activations with FP32 accumulation, predicts the tree-order surface potential
 * and its normal derivative; the prediction replaces the zero initial guess
 * that the reference passes to DGMRES (xvct = 0, src/main.f90:115).  GMRES
 * still iterates to the same tolerance, and falls back to the zero start when
 * the guess's preconditioned residual exceeds that of x0 = 0.
 *
 * Packed parameters (tabi_ws_param_count() floats): feature means[12], feature
 * scales[12], output scales[2], 6 unused, W1[128][12], b1[128], W2[128][128],
 * b2[128], W3[128][128], b3[128], W4[2][128], b4[2] (row-major, out x in).
 * Features per face (unnormalised, as tabi_ctx_ws_features returns them),
 * with s = 1/rms(b_1..b_n): s*b_i and s*b_{n+i} (the right-hand side: Coulomb
 * potential and normal field of all charges); the same potential and normal
 * field restricted by Gaussian windows of 5 A and 10 A around the face (4);
 * area; curvature (leaf-local normal divergence); distances to the nearest
 * and second-nearest charge; normal . direction from the nearest charge;
 * a smooth count of charges within 10 A.  Targets for training are the
 * tree-order solution scaled by s: (s*phi_i, s*dphi_i/dn). */
typedef struct tabi_ws tabi_ws;
int64_t tabi_ws_param_count(void);
int tabi_ws_create(const float *params, int64_t count, tabi_ws **out);
void tabi_ws_destroy(tabi_ws *m);
int tabi_ws_get_params(tabi_ws *m, float *out, int64_t count);
/* features feat_out[n][12] (tree order) of a context's surface for nq charges;
 * scale_out receives s */
int tabi_ctx_ws_features(tabi_ctx *c, int32_t nq, const double *qpos, const double *q,
                         double eps_p, float *feat_out, double *scale_out);
/* x0_out[2n] = predicted tree-order guess (phi then dphi/dn) from feat[n][12] */
int tabi_ws_predict(tabi_ws *m, int64_t n, const float *feat, double scale, double *x0_out);
/* gradient of the loss (1/2ns) sum |out - target/out_scale|^2 over ns samples
 * (packed layout; the first 32 entries are zero) and the loss */
int tabi_ws_grad(tabi_ws *m, int64_t ns, const float *feat, const float *target,
                 float *grad_out, double *loss_out);
/* Adam training on the GPU: epochs over ns samples in shuffled mini-batches
 * (batch <= 131072), cosine-decayed learning rate from lr to lr/20;
 * epoch_loss[epochs] (may be NULL) receives the mean loss of each epoch */
int tabi_ws_train(tabi_ws *m, int64_t ns, const float *feat, const float *target,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss);
/* GMRES (as tabi_ctx_gmres) from the model's guess for the charges; x gets
 * the solution, ws_used 1 if the guess was kept, 0 if the zero start was used */
int tabi_ctx_gmres_ws(tabi_ctx *c, tabi_ws *model, int32_t nq, const double *qpos,
                      const double *q, double eps_p, const double *b, double *x, double tol,
                      int32_t maxl, int32_t nrmax,
                      int32_t *iter, double *err, int32_t *ierr, int32_t *ws_used);
/* tabi_solve with a learned initial guess (model may be NULL: zero start) */
int tabi_solve_ws(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
                  const double *area, int32_t nq, const double *qpos, const double *q,
                  double eps_p, double tol, tabi_ws *model, double *x_out,
                  int32_t *orderarr_out, int32_t *iter, double *err, int32_t *ierr,
                  double *soleng, double *couleng, int32_t *ws_used);

/* tabi_solve_ws with every array already resident in device memory: d_xyz,
 * d_nrm, d_area, d_qpos, d_q and d_x_out (may be NULL) are device pointers
 * (original face order in, tree order out); orderarr_out stays a host array. */
int tabi_solve_dev(const tabi_params *prm, int32_t n, const double *d_xyz, const double *d_nrm,
                   const double *d_area, int32_t nq, const double *d_qpos, const double *d_q,
                   double eps_p, double tol, tabi_ws *model, double *d_x_out,
                   int32_t *orderarr_out, int32_t *iter, double *err, int32_t *ierr,
                   double *soleng, double *couleng, int32_t *ws_used);

#ifdef __cplusplus
}
#endif
#endif
