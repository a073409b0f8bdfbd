/* CPU oracle for the TABI-PB treecode-accelerated GMRES solve.
 *
 * TEST INFRASTRUCTURE ONLY (see tabi_oracle.h).  Built with -ffp-contract=off so
 * that each expression rounds the way the reference's Fortran (x86-64, no FMA)
 * rounds it; expressions keep the reference's operand order on purpose.
 */
#include "tabi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.141592653589793238462643383279; /* acos(-1d0), src/main.f90:60 */
#define ONE_OVER_4PI (0.25 / PI)                           /* src/main.f90:61 */

double or_kappa(double bulk_strength, double eps_w)
{
    /* src/main.f90:62: the literal 8.430325455 has default (single) kind */
    double kappa2 = (double)8.430325455f * bulk_strength / eps_w;
    return sqrt(kappa2);
}

/* ------------------------------------------------------------------------- */
/* geometry, culling, right-hand side                                         */
/* ------------------------------------------------------------------------- */
static double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static double heron(double aa, double bb, double cc)
{
    /* triangle_area, src/readin.f90:336-340 */
    double s = 0.5 * (aa + bb + cc);
    return sqrt(s * (s - aa) * (s - bb) * (s - cc));
}

static double edge(const double *p, const double *q)
{
    double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return sqrt(dot3(d, d));
}

void or_face_geometry(int32_t nface, const double *vpos, const double *vnrm,
                      const int32_t *faces, double *xyz, double *nrm, double *area)
{
    const double third = 1.0 / 3.0;
    for (int32_t i = 0; i < nface; ++i) {
        double r0[3] = {0, 0, 0}, v0[3] = {0, 0, 0};
        const double *r[3];
        for (int k = 0; k < 3; ++k) {               /* src/main.f90:572-577 */
            const double *p = vpos + 3 * (int64_t)faces[3 * i + k];
            const double *v = vnrm + 3 * (int64_t)faces[3 * i + k];
            for (int d = 0; d < 3; ++d) {
                r0[d] = r0[d] + third * p[d];
                v0[d] = v0[d] + third * v[d];
            }
            r[k] = p;
        }
        double vn = sqrt(dot3(v0, v0));              /* src/main.f90:581 */
        for (int d = 0; d < 3; ++d) {
            xyz[3 * i + d] = r0[d];
            nrm[3 * i + d] = v0[d] / vn;
        }
        area[i] = heron(edge(r[0], r[1]), edge(r[0], r[2]), edge(r[1], r[2]));
    }
}

static void centroid(const double *vpos, const int32_t *f, double *c)
{
    const double third = 1.0 / 3.0;
    c[0] = c[1] = c[2] = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) c[d] = c[d] + third * vpos[3 * (int64_t)f[k] + d];
}

int32_t or_cull_faces(int32_t nface, const double *vpos, const int32_t *faces,
                      int32_t *faces_out, double *s_area)
{
    /* src/readin.f90:277-313: neighbours are the 10 preceding faces in the
     * original numbering, including faces that were themselves deleted */
    double sa = 0.0;
    int32_t kept = 0;
    for (int32_t i = 0; i < nface; ++i) {
        const int32_t *f = faces + 3 * (int64_t)i;
        double xx[3];
        centroid(vpos, f, xx);
        const double *p0 = vpos + 3 * (int64_t)f[0], *p1 = vpos + 3 * (int64_t)f[1],
                     *p2 = vpos + 3 * (int64_t)f[2];
        double a = heron(edge(p0, p1), edge(p0, p2), edge(p1, p2));
        int alert = 0;
        for (int32_t ii = (i - 10 > 0 ? i - 10 : 0); ii < i; ++ii) {
            double yy[3], d[3];
            centroid(vpos, faces + 3 * (int64_t)ii, yy);
            for (int k = 0; k < 3; ++k) d[k] = xx[k] - yy[k];
            if (dot3(d, d) < 1.0e-5) alert = 1;
        }
        if (!(a < 1.0e-5 || alert)) {
            faces_out[3 * kept + 0] = f[0];
            faces_out[3 * kept + 1] = f[1];
            faces_out[3 * kept + 2] = f[2];
            ++kept;
        }
        sa = sa + a;
    }
    if (s_area) *s_area = sa;
    return kept;
}

void or_rhs(int32_t n, const double *xyz, const double *nrm, int32_t nq,
            const double *qpos, const double *q, double eps_p, double *b)
{
    for (int32_t i = 0; i < n; ++i) {                /* src/main.f90:597-607 */
        const double *r0 = xyz + 3 * (int64_t)i, *v0 = nrm + 3 * (int64_t)i;
        double b1 = 0.0, b2 = 0.0;
        for (int32_t j = 0; j < nq; ++j) {
            const double *rr = qpos + 3 * (int64_t)j;
            double d[3] = {rr[0] - r0[0], rr[1] - r0[1], rr[2] - r0[2]};
            double rs = sqrt(dot3(d, d));
            double g0 = 1.0 / (4.0 * PI * rs);
            double ct = dot3(v0, d) / rs;
            double g1 = ct / (rs * rs) / 4.0 / PI;
            b1 = b1 + q[j] * g0 / eps_p;
            b2 = b2 + q[j] * g1 / eps_p;
        }
        b[i] = b1;
        b[n + i] = b2;
    }
}

/* ------------------------------------------------------------------------- */
/* tree build                                                                 */
/* ------------------------------------------------------------------------- */
/* PARTITION (src/treecode3d_pb.f:1677-1759): hole-walk partition of
 * a[ibeg..iend] (0-based, inclusive) about val, carrying b, c and ind.  Returns
 * midind with a[ibeg..midind] <= val < a[midind+1..iend]. */
static int32_t partition(double *a, double *b, double *c, int32_t *ind,
                         int32_t ibeg, int32_t iend, double val)
{
    if (ibeg < iend) {
        double ta = a[ibeg], tb = b[ibeg], tc = c[ibeg];
        int32_t tind = ind[ibeg];
        a[ibeg] = val;
        int32_t upper = ibeg, lower = iend;
        while (upper != lower) {
            while (upper < lower && val < a[lower]) --lower;
            if (upper != lower) {
                a[upper] = a[lower]; b[upper] = b[lower]; c[upper] = c[lower];
                ind[upper] = ind[lower];
            }
            while (upper < lower && val >= a[upper]) ++upper;
            if (upper != lower) {
                a[lower] = a[upper]; b[lower] = b[upper]; c[lower] = c[upper];
                ind[lower] = ind[upper];
            }
        }
        int32_t midind = upper;
        if (ta > val) midind = upper - 1;
        a[upper] = ta; b[upper] = tb; c[upper] = tc; ind[upper] = tind;
        return midind;
    } else if (ibeg == iend) {
        return a[ibeg] <= val ? ibeg : ibeg - 1;
    }
    return ibeg - 1;
}

typedef struct {
    double *x, *y, *z;
    int32_t *ord;
    int32_t maxpar;
    or_node *nodes;
    int32_t nnode;
} tree_ctx;

static void minmax(const double *a, int32_t i0, int32_t i1, double *mn, double *mx)
{
    double lo = a[i0], hi = a[i0];
    for (int32_t i = i0 + 1; i <= i1; ++i) {
        if (a[i] < lo) lo = a[i];
        if (a[i] > hi) hi = a[i];
    }
    *mn = lo; *mx = hi;
}

/* CREATE_TREE (src/treecode3d_pb.f:142-281) with PARTITION_8 (284-356) inlined */
static int32_t create_tree(tree_ctx *t, int32_t ibeg, int32_t iend, const double bx[6],
                           int32_t level)
{
    int32_t id = t->nnode++;
    or_node *p = &t->nodes[id];
    memset(p, 0, sizeof(*p));
    p->numpar = iend - ibeg + 1;
    p->xmin = bx[0]; p->xmax = bx[1]; p->ymin = bx[2];
    p->ymax = bx[3]; p->zmin = bx[4]; p->zmax = bx[5];
    double xl = p->xmax - p->xmin, yl = p->ymax - p->ymin, zl = p->zmax - p->zmin;
    double lmax = xl;
    if (yl > lmax) lmax = yl;
    if (zl > lmax) lmax = zl;
    p->xmid = (p->xmax + p->xmin) / 2.0;
    p->ymid = (p->ymax + p->ymin) / 2.0;
    p->zmid = (p->zmax + p->zmin) / 2.0;
    double t1 = p->xmax - p->xmid, t2 = p->ymax - p->ymid, t3 = p->zmax - p->zmid;
    p->radius = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    p->ibeg = ibeg; p->iend = iend; p->level = level; p->nchild = 0;
    if (p->numpar <= t->maxpar) return id;

    int32_t ind[8][2];
    double bxs[8][6];
    memset(ind, 0, sizeof(ind));
    for (int k = 0; k < 6; ++k) bxs[0][k] = bx[k];
    ind[0][0] = ibeg; ind[0][1] = iend;
    double xm = p->xmid, ym = p->ymid, zm = p->zmid;
    double critlen = lmax / sqrt(2.0);
    int nch = 1;
    if (xl >= critlen) {
        int32_t mid = partition(t->x, t->y, t->z, t->ord, ind[0][0], ind[0][1], xm);
        ind[1][0] = mid + 1; ind[1][1] = ind[0][1]; ind[0][1] = mid;
        memcpy(bxs[1], bxs[0], sizeof(bxs[0]));
        bxs[0][1] = xm; bxs[1][0] = xm;
        nch *= 2;
    }
    if (yl >= critlen) {
        for (int i = 0; i < nch; ++i) {
            int32_t mid = partition(t->y, t->x, t->z, t->ord, ind[i][0], ind[i][1], ym);
            ind[nch + i][0] = mid + 1; ind[nch + i][1] = ind[i][1]; ind[i][1] = mid;
            memcpy(bxs[nch + i], bxs[i], sizeof(bxs[0]));
            bxs[i][3] = ym; bxs[nch + i][2] = ym;
        }
        nch *= 2;
    }
    if (zl >= critlen) {
        for (int i = 0; i < nch; ++i) {
            int32_t mid = partition(t->z, t->x, t->y, t->ord, ind[i][0], ind[i][1], zm);
            ind[nch + i][0] = mid + 1; ind[nch + i][1] = ind[i][1]; ind[i][1] = mid;
            memcpy(bxs[nch + i], bxs[i], sizeof(bxs[0]));
            bxs[i][5] = zm; bxs[nch + i][4] = zm;
        }
        nch *= 2;
    }
    /* shrink boxes holding at least two particles (src/treecode3d_pb.f:249-258) */
    for (int i = 0; i < 8; ++i) {
        if (ind[i][0] < ind[i][1]) {
            minmax(t->x, ind[i][0], ind[i][1], &bxs[i][0], &bxs[i][1]);
            minmax(t->y, ind[i][0], ind[i][1], &bxs[i][2], &bxs[i][3]);
            minmax(t->z, ind[i][0], ind[i][1], &bxs[i][4], &bxs[i][5]);
        }
    }
    for (int i = 0; i < nch; ++i) {
        if (ind[i][0] <= ind[i][1]) {
            int32_t c = create_tree(t, ind[i][0], ind[i][1], bxs[i], level + 1);
            p = &t->nodes[id];
            p->child[p->nchild++] = c;
        }
    }
    return id;
}

int32_t or_build_tree(int32_t n, double *xyz, int32_t maxparnode, int32_t *orderarr,
                      or_node *nodes)
{
    tree_ctx t;
    t.x = malloc(sizeof(double) * n);
    t.y = malloc(sizeof(double) * n);
    t.z = malloc(sizeof(double) * n);
    t.ord = orderarr;
    t.maxpar = maxparnode;
    t.nodes = nodes;
    t.nnode = 0;
    for (int32_t i = 0; i < n; ++i) {
        t.x[i] = xyz[3 * i]; t.y[i] = xyz[3 * i + 1]; t.z[i] = xyz[3 * i + 2];
        orderarr[i] = i;
    }
    double bx[6];                                    /* SETUP, tc.f:120-125 */
    minmax(t.x, 0, n - 1, &bx[0], &bx[1]);
    minmax(t.y, 0, n - 1, &bx[2], &bx[3]);
    minmax(t.z, 0, n - 1, &bx[4], &bx[5]);
    create_tree(&t, 0, n - 1, bx, 0);
    for (int32_t i = 0; i < n; ++i) {
        xyz[3 * i] = t.x[i]; xyz[3 * i + 1] = t.y[i]; xyz[3 * i + 2] = t.z[i];
    }
    free(t.x); free(t.y); free(t.z);
    return t.nnode;
}

static void leaves_rec(const or_node *nodes, int32_t id, int32_t *ib, int32_t *ie, int32_t *k)
{
    const or_node *p = &nodes[id];
    if (p->nchild == 0) {
        if (ib) { ib[*k] = p->ibeg; ie[*k] = p->iend; }
        ++*k;
        return;
    }
    for (int c = 0; c < p->nchild; ++c) leaves_rec(nodes, p->child[c], ib, ie, k);
}

int32_t or_leaves(const or_node *nodes, int32_t *ibeg, int32_t *iend)
{
    int32_t k = 0;
    leaves_rec(nodes, 0, ibeg, iend, &k);
    return k;
}

/* ------------------------------------------------------------------------- */
/* Taylor coefficients (COMP_TCOEFF, src/treecode3d_pb.f:684-855)             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int tl;        /* torderlim */
    int dim;       /* tl + 3 (indices -2..tl) */
    double *a, *b;
    double *cf1, *cf2, *cf3;
} tco;

#define IDX(T, i, j, k) ((((i) + 2) * (T)->dim + ((j) + 2)) * (T)->dim + ((k) + 2))
#define A_(i, j, k) T->a[IDX(T, i, j, k)]
#define B_(i, j, k) T->b[IDX(T, i, j, k)]

static void tco_init(tco *T, int tl)
{
    T->tl = tl;
    T->dim = tl + 3;
    size_t n3 = (size_t)T->dim * T->dim * T->dim;
    T->a = calloc(n3, sizeof(double));
    T->b = calloc(n3, sizeof(double));
    T->cf1 = calloc(tl + 1, sizeof(double));
    T->cf2 = calloc(tl + 1, sizeof(double));
    T->cf3 = calloc(tl + 1, sizeof(double));
    for (int i = 1; i <= tl; ++i) {                  /* src/treecode3d_pb.f:111-116 */
        double t1 = 1.0 / (double)i;
        T->cf1[i] = t1;
        T->cf2[i] = 1.0 - 0.5 * t1;
        T->cf3[i] = 1.0 - t1;
    }
}

static void tco_free(tco *T)
{
    free(T->a); free(T->b); free(T->cf1); free(T->cf2); free(T->cf3);
}

static void tco_compute(tco *T, double dx, double dy, double dz, double kappa)
{
    const int tl = T->tl;
    double c1[64];
    const double *cf2 = T->cf2, *cf3 = T->cf3;
    for (int i = 1; i <= tl; ++i) c1[i] = T->cf1[i] * kappa;
    double ddx = 2.0 * dx, ddy = 2.0 * dy, ddz = 2.0 * dz;
    double kx = kappa * dx, ky = kappa * dy, kz = kappa * dz;
    double dist = dx * dx + dy * dy + dz * dz;
    double fac = 1.0 / dist;
    dist = sqrt(dist);

    B_(0, 0, 0) = exp(-kappa * dist);
    A_(0, 0, 0) = B_(0, 0, 0) / dist;

    B_(1, 0, 0) = kx * A_(0, 0, 0);
    B_(0, 1, 0) = ky * A_(0, 0, 0);
    B_(0, 0, 1) = kz * A_(0, 0, 0);
    A_(1, 0, 0) = fac * dx * (A_(0, 0, 0) + kappa * B_(0, 0, 0));
    A_(0, 1, 0) = fac * dy * (A_(0, 0, 0) + kappa * B_(0, 0, 0));
    A_(0, 0, 1) = fac * dz * (A_(0, 0, 0) + kappa * B_(0, 0, 0));

    for (int i = 2; i <= tl; ++i) {
        B_(i, 0, 0) = c1[i] * (dx * A_(i - 1, 0, 0) - A_(i - 2, 0, 0));
        B_(0, i, 0) = c1[i] * (dy * A_(0, i - 1, 0) - A_(0, i - 2, 0));
        B_(0, 0, i) = c1[i] * (dz * A_(0, 0, i - 1) - A_(0, 0, i - 2));
        A_(i, 0, 0) = fac * (ddx * cf2[i] * A_(i - 1, 0, 0) - cf3[i] * A_(i - 2, 0, 0) +
                             c1[i] * (dx * B_(i - 1, 0, 0) - B_(i - 2, 0, 0)));
        A_(0, i, 0) = fac * (ddy * cf2[i] * A_(0, i - 1, 0) - cf3[i] * A_(0, i - 2, 0) +
                             c1[i] * (dy * B_(0, i - 1, 0) - B_(0, i - 2, 0)));
        A_(0, 0, i) = fac * (ddz * cf2[i] * A_(0, 0, i - 1) - cf3[i] * A_(0, 0, i - 2) +
                             c1[i] * (dz * B_(0, 0, i - 1) - B_(0, 0, i - 2)));
    }

    B_(1, 1, 0) = kx * A_(0, 1, 0);
    B_(1, 0, 1) = kx * A_(0, 0, 1);
    B_(0, 1, 1) = ky * A_(0, 0, 1);
    A_(1, 1, 0) = fac * (dx * A_(0, 1, 0) + ddy * A_(1, 0, 0) + kx * B_(0, 1, 0));
    A_(1, 0, 1) = fac * (dx * A_(0, 0, 1) + ddz * A_(1, 0, 0) + kx * B_(0, 0, 1));
    A_(0, 1, 1) = fac * (dy * A_(0, 0, 1) + ddz * A_(0, 1, 0) + ky * B_(0, 0, 1));

    for (int i = 2; i <= tl - 1; ++i) {
        B_(1, 0, i) = kx * A_(0, 0, i);
        B_(0, 1, i) = ky * A_(0, 0, i);
        B_(0, i, 1) = kz * A_(0, i, 0);
        B_(1, i, 0) = kx * A_(0, i, 0);
        B_(i, 1, 0) = ky * A_(i, 0, 0);
        B_(i, 0, 1) = kz * A_(i, 0, 0);
        A_(1, 0, i) = fac * (dx * A_(0, 0, i) + ddz * A_(1, 0, i - 1) - A_(1, 0, i - 2) +
                             kx * B_(0, 0, i));
        A_(0, 1, i) = fac * (dy * A_(0, 0, i) + ddz * A_(0, 1, i - 1) - A_(0, 1, i - 2) +
                             ky * B_(0, 0, i));
        A_(0, i, 1) = fac * (dz * A_(0, i, 0) + ddy * A_(0, i - 1, 1) - A_(0, i - 2, 1) +
                             kz * B_(0, i, 0));
        A_(1, i, 0) = fac * (dx * A_(0, i, 0) + ddy * A_(1, i - 1, 0) - A_(1, i - 2, 0) +
                             kx * B_(0, i, 0));
        A_(i, 1, 0) = fac * (dy * A_(i, 0, 0) + ddx * A_(i - 1, 1, 0) - A_(i - 2, 1, 0) +
                             ky * B_(i, 0, 0));
        A_(i, 0, 1) = fac * (dz * A_(i, 0, 0) + ddx * A_(i - 1, 0, 1) - A_(i - 2, 0, 1) +
                             kz * B_(i, 0, 0));
    }

    for (int i = 2; i <= tl - 2; ++i) {
        for (int j = 2; j <= tl - i; ++j) {
            B_(i, j, 0) = c1[i] * (dx * A_(i - 1, j, 0) - A_(i - 2, j, 0));
            B_(i, 0, j) = c1[i] * (dx * A_(i - 1, 0, j) - A_(i - 2, 0, j));
            B_(0, i, j) = c1[i] * (dy * A_(0, i - 1, j) - A_(0, i - 2, j));
            A_(i, j, 0) = fac * (ddx * cf2[i] * A_(i - 1, j, 0) + ddy * A_(i, j - 1, 0) -
                                 cf3[i] * A_(i - 2, j, 0) - A_(i, j - 2, 0) +
                                 c1[i] * (dx * B_(i - 1, j, 0) - B_(i - 2, j, 0)));
            A_(i, 0, j) = fac * (ddx * cf2[i] * A_(i - 1, 0, j) + ddz * A_(i, 0, j - 1) -
                                 cf3[i] * A_(i - 2, 0, j) - A_(i, 0, j - 2) +
                                 c1[i] * (dx * B_(i - 1, 0, j) - B_(i - 2, 0, j)));
            A_(0, i, j) = fac * (ddy * cf2[i] * A_(0, i - 1, j) + ddz * A_(0, i, j - 1) -
                                 cf3[i] * A_(0, i - 2, j) - A_(0, i, j - 2) +
                                 c1[i] * (dy * B_(0, i - 1, j) - B_(0, i - 2, j)));
        }
    }

    B_(1, 1, 1) = kx * A_(0, 1, 1);
    A_(1, 1, 1) = fac * (dx * A_(0, 1, 1) + ddy * A_(1, 0, 1) + ddz * A_(1, 1, 0) +
                         kx * B_(0, 1, 1));

    for (int i = 2; i <= tl - 2; ++i) {
        B_(1, 1, i) = kx * A_(0, 1, i);
        B_(1, i, 1) = kx * A_(0, i, 1);
        B_(i, 1, 1) = ky * A_(i, 0, 1);
        A_(1, 1, i) = fac * (dx * A_(0, 1, i) + ddy * A_(1, 0, i) + ddz * A_(1, 1, i - 1) -
                             A_(1, 1, i - 2) + kx * B_(0, 1, i));
        A_(1, i, 1) = fac * (dx * A_(0, i, 1) + ddy * A_(1, i - 1, 1) + ddz * A_(1, i, 0) -
                             A_(1, i - 2, 1) + kx * B_(0, i, 1));
        A_(i, 1, 1) = fac * (dy * A_(i, 0, 1) + ddx * A_(i - 1, 1, 1) + ddz * A_(i, 1, 0) -
                             A_(i - 2, 1, 1) + ky * B_(i, 0, 1));
    }

    for (int i = 2; i <= tl - 3; ++i) {
        for (int j = 2; j <= tl - i; ++j) {
            B_(1, i, j) = kx * A_(0, i, j);
            B_(i, 1, j) = ky * A_(i, 0, j);
            B_(i, j, 1) = kz * A_(i, j, 0);
            A_(1, i, j) = fac * (dx * A_(0, i, j) + ddy * A_(1, i - 1, j) + ddz * A_(1, i, j - 1) -
                                 A_(1, i - 2, j) - A_(1, i, j - 2) + kx * B_(0, i, j));
            A_(i, 1, j) = fac * (dy * A_(i, 0, j) + ddx * A_(i - 1, 1, j) + ddz * A_(i, 1, j - 1) -
                                 A_(i - 2, 1, j) - A_(i, 1, j - 2) + ky * B_(i, 0, j));
            A_(i, j, 1) = fac * (dz * A_(i, j, 0) + ddx * A_(i - 1, j, 1) + ddy * A_(i, j - 1, 1) -
                                 A_(i - 2, j, 1) - A_(i, j - 2, 1) + kz * B_(i, j, 0));
        }
    }

    for (int k = 2; k <= tl - 4; ++k) {
        for (int j = 2; j <= tl - 2 - k; ++j) {
            for (int i = 2; i <= tl - k - j; ++i) {
                B_(i, j, k) = c1[i] * (dx * A_(i - 1, j, k) - A_(i - 2, j, k));
                A_(i, j, k) = fac * (ddx * cf2[i] * A_(i - 1, j, k) + ddy * A_(i, j - 1, k) +
                                     ddz * A_(i, j, k - 1) - cf3[i] * A_(i - 2, j, k) -
                                     A_(i, j - 2, k) - A_(i, j, k - 2) +
                                     c1[i] * (dx * B_(i - 1, j, k) - B_(i - 2, j, k)));
            }
        }
    }
}

void or_taylor_coeffs(double dx, double dy, double dz, double kappa, int32_t lim, double *out)
{
    tco T;
    tco_init(&T, lim);
    tco *TT = &T;
    tco_compute(TT, dx, dy, dz, kappa);
    for (int i = 0; i <= lim; ++i)
        for (int j = 0; j <= lim; ++j)
            for (int k = 0; k <= lim; ++k) {
                double v = 0.0;
                if (i + j + k <= lim) v = TT->a[IDX(TT, i, j, k)];
                out[(i * (lim + 1) + j) * (lim + 1) + k] = v;
            }
    tco_free(&T);
}

/* ------------------------------------------------------------------------- */
/* treecode matvec                                                            */
/* ------------------------------------------------------------------------- */
/* derivative offsets kk and coefficients der_cof (src/main.f90:358-399) */
static const int KK[16][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0},
                              {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 0}, {0, 2, 0},
                              {0, 1, 1}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};

typedef struct {
    const or_params *prm;
    int32_t n;
    const double *xyz, *nrm, *area, *x;
    const or_node *nodes;
    double *sq;          /* per-particle source charges q(i,m), m=0..15 */
    double **ms;         /* per-node moments, lazily formed; [m][i][j][k] */
    int p, mdim;         /* order, p+1 */
    double *dercof;      /* [m][i][j][k] */
    tco T;
    double tx, ty, tz;   /* target position */
    double tq[16];       /* target charges */
    int64_t npair_near, npair_far, nmom;
    int32_t self;        /* target index (excluded) */
} mv_ctx;

#define MIDX(c, m, i, j, k) ((((m) * (c)->mdim + (i)) * (c)->mdim + (j)) * (c)->mdim + (k))

static void comp_ms(mv_ctx *c, int32_t id)
{
    /* COMP_MS, src/treecode3d_pb.f:859-916: moments for m = 1..7, 8, 11, 14 */
    const or_node *nd = &c->nodes[id];
    int p = c->p;
    double *ms = calloc((size_t)16 * c->mdim * c->mdim * c->mdim, sizeof(double));
    for (int32_t i = nd->ibeg; i <= nd->iend; ++i) {
        double dx = c->xyz[3 * i] - nd->xmid;
        double dy = c->xyz[3 * i + 1] - nd->ymid;
        double dz = c->xyz[3 * i + 2] - nd->zmid;
        const double *q = c->sq + 16 * (int64_t)i;
        double tz = 1.0;
        for (int k3 = 0; k3 <= p; ++k3) {
            double ty = 1.0;
            for (int k2 = 0; k2 <= p - k3; ++k2) {
                double tx = 1.0;
                for (int k1 = 0; k1 <= p - k3 - k2; ++k1) {
                    double txyz = tx * ty * tz;
                    for (int m = 0; m < 7; ++m)
                        ms[MIDX(c, m, k1, k2, k3)] = ms[MIDX(c, m, k1, k2, k3)] + q[m] * txyz;
                    for (int m = 7; m < 16; m += 3)
                        ms[MIDX(c, m, k1, k2, k3)] = ms[MIDX(c, m, k1, k2, k3)] + q[m] * txyz;
                    tx = tx * dx;
                }
                ty = ty * dy;
            }
            tz = tz * dz;
        }
    }
    size_t blk = (size_t)c->mdim * c->mdim * c->mdim;
    memcpy(ms + 8 * blk, ms + 7 * blk, blk * sizeof(double));
    memcpy(ms + 9 * blk, ms + 7 * blk, blk * sizeof(double));
    memcpy(ms + 11 * blk, ms + 10 * blk, blk * sizeof(double));
    memcpy(ms + 12 * blk, ms + 10 * blk, blk * sizeof(double));
    memcpy(ms + 14 * blk, ms + 13 * blk, blk * sizeof(double));
    memcpy(ms + 15 * blk, ms + 13 * blk, blk * sizeof(double));
    c->ms[id] = ms;
    c->nmom += nd->numpar;
}

static void far_pb(mv_ctx *c, int32_t id, double peng[2])
{
    /* COMPP_TREE_PB, src/treecode3d_pb.f:439-486 */
    const or_node *nd = &c->nodes[id];
    const double *ms = c->ms[id];
    double kapa[2] = {0.0, c->prm->kappa};
    double pt[16][2];
    int p = c->p;
    tco *T = &c->T;
    for (int ikp = 0; ikp < 2; ++ikp) {
        tco_compute(T, c->tx - nd->xmid, c->ty - nd->ymid, c->tz - nd->zmid, kapa[ikp]);
        for (int m = 0; m < 16; ++m) {
            double s = 0.0;
            for (int k = 0; k <= p; ++k)
                for (int j = 0; j <= p - k; ++j)
                    for (int i = 0; i <= p - k - j; ++i)
                        s = s + c->dercof[MIDX(c, m, i, j, k)] *
                                    A_(i + KK[m][0], j + KK[m][1], k + KK[m][2]) *
                                    ms[MIDX(c, m, i, j, k)];
            pt[m][ikp] = c->tq[m] * s;
        }
    }
    double eps = c->prm->eps;
    double s24_1 = pt[1][0] + pt[2][0] + pt[3][0], s24_2 = pt[1][1] + pt[2][1] + pt[3][1];
    double s57_1 = pt[4][0] + pt[5][0] + pt[6][0], s57_2 = pt[4][1] + pt[5][1] + pt[6][1];
    double s8_1 = 0.0, s8_2 = 0.0;
    for (int m = 7; m < 16; ++m) { s8_1 = s8_1 + pt[m][0]; s8_2 = s8_2 + pt[m][1]; }
    double sl1 = pt[0][0] - pt[0][1];
    double sl2 = eps * s24_2 - s24_1;
    double sl3 = -(s57_1 - 1.0 / eps * s57_2);
    double sl4 = s8_2 - s8_1;
    peng[0] = sl1 + sl2;
    peng[1] = sl3 + sl4;
}

/* one source-target pair of the four BIE kernels (COMPP_DIRECT_PB,
 * src/treecode3d_pb.f:965-1001); adds area-weighted contributions to peng */
static void pair_pb(double kappa, double eps, const double *s, const double *v0,
                    const double *r, const double *v, double area, double x1, double x2,
                    double peng[2])
{
    double d[3] = {r[0] - s[0], r[1] - s[1], r[2] - s[2]};
    double rs = sqrt(dot3(d, d));
    double g0 = ONE_OVER_4PI / rs;
    double krs = kappa * rs;
    double ek = exp(-krs);
    double gk = ek * g0;
    double ct = dot3(v, d) / rs;
    double ct0 = dot3(v0, d) / rs;
    double tp1 = g0 / rs;
    double tp2 = (1.0 + krs) * ek;
    double g10 = ct0 * tp1, g20 = tp2 * g10;
    double g1 = ct * tp1, g2 = tp2 * g1;
    double g3 = (dot3(v0, v) - 3.0 * ct0 * ct) / rs * tp1;
    double g4 = tp2 * g3 - kappa * kappa * ct0 * ct * gk;
    double l1 = g1 - eps * g2, l2 = g0 - gk, l3 = g4 - g3, l4 = g10 - g20 / eps;
    peng[0] = peng[0] + (l1 * x1 + l2 * x2) * area;
    peng[1] = peng[1] + (l3 * x1 + l4 * x2) * area;
}

static void direct_pb(mv_ctx *c, int32_t ibeg, int32_t iend, double peng[2])
{
    double s[3] = {c->tx, c->ty, c->tz};
    const double *v0 = c->nrm + 3 * (int64_t)c->self;
    peng[0] = peng[1] = 0.0;
    for (int32_t j = ibeg; j <= iend; ++j) {
        if (j == c->self) continue; /* zero-area self term, src/treecode3d_pb.f:1169-1172 */
        pair_pb(c->prm->kappa, c->prm->eps, s, v0, c->xyz + 3 * (int64_t)j,
                c->nrm + 3 * (int64_t)j, c->area[j], c->x[j], c->x[c->n + j], peng);
        c->npair_near++;
    }
}

static void compp_tree(mv_ctx *c, int32_t id, double peng[2])
{
    /* COMPP_TREE, src/treecode3d_pb.f:359-436 */
    const or_node *nd = &c->nodes[id];
    double tx = nd->xmid - c->tx, ty = nd->ymid - c->ty, tz = nd->zmid - c->tz;
    double dist = sqrt(tx * tx + ty * ty + tz * tz);
    peng[0] = peng[1] = 0.0;
    if (nd->radius < dist * c->prm->theta && nd->numpar > 40) {
        if (!c->ms[id]) comp_ms(c, id);
        far_pb(c, id, peng);
        c->npair_far++;
    } else if (nd->nchild == 0) {
        direct_pb(c, nd->ibeg, nd->iend, peng);
    } else {
        for (int k = 0; k < nd->nchild; ++k) {
            double pl[2];
            compp_tree(c, nd->child[k], pl);
            peng[0] = peng[0] + pl[0];
            peng[1] = peng[1] + pl[1];
        }
    }
}

static int32_t count_nodes(const or_node *nodes, int32_t id)
{
    int32_t k = 1;
    for (int c = 0; c < nodes[id].nchild; ++c) k += count_nodes(nodes, nodes[id].child[c]);
    return k;
}

static void treecode_targets(const or_params *prm, int32_t n, const double *xyz,
                             const double *nrm, const double *area, const or_node *nodes,
                             const double *x, int32_t nt, const int32_t *targets, double *y,
                             int64_t *counts);

void or_treecode_matvec(const or_params *prm, int32_t n, const double *xyz,
                        const double *nrm, const double *area, const or_node *nodes,
                        const double *x, double *y, int64_t *counts)
{
    treecode_targets(prm, n, xyz, nrm, area, nodes, x, -1, NULL, y, counts);
}

void or_treecode_matvec_targets(const or_params *prm, int32_t n, const double *xyz,
                                const double *nrm, const double *area, const or_node *nodes,
                                const double *x, int32_t nt, const int32_t *targets,
                                double *y, int64_t *counts)
{
    treecode_targets(prm, n, xyz, nrm, area, nodes, x, nt, targets, y, counts);
}

static void treecode_targets(const or_params *prm, int32_t n, const double *xyz,
                             const double *nrm, const double *area, const or_node *nodes,
                             const double *x, int32_t nt, const int32_t *targets, double *y,
                             int64_t *counts)
{
    mv_ctx c;
    memset(&c, 0, sizeof(c));
    c.prm = prm; c.n = n; c.xyz = xyz; c.nrm = nrm; c.area = area; c.x = x; c.nodes = nodes;
    c.p = prm->order;
    c.mdim = c.p + 1;
    int32_t nn = count_nodes(nodes, 0);
    c.ms = calloc(nn, sizeof(double *));
    tco_init(&c.T, prm->order + 3);   /* torderlim = p+2+orderoffset, tc.f:73-81 */
    /* der_cof (src/main.f90:380-399) */
    c.dercof = malloc(sizeof(double) * 16 * c.mdim * c.mdim * c.mdim);
    for (int m = 0; m < 16; ++m)
        for (int i = 0; i <= c.p; ++i)
            for (int j = 0; j <= c.p; ++j)
                for (int k = 0; k <= c.p; ++k) {
                    double v = 1.0;
                    int ijk[3] = {i, j, k};
                    if (i + j + k <= c.p)
                        for (int d = 0; d < 3; ++d)
                            for (int t = 1; t <= KK[m][d]; ++t) v = v * (double)(ijk[d] + t);
                    c.dercof[MIDX(&c, m, i, j, k)] = v * ONE_OVER_4PI;
                }
    /* pb_kernel source charges (src/main.f90:671-692) */
    c.sq = malloc(sizeof(double) * 16 * (size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        double *q = c.sq + 16 * (int64_t)i;
        const double *nv = nrm + 3 * (int64_t)i;
        double phi = x[i], psi = x[n + i], a = area[i];
        q[0] = a * psi;
        for (int d = 0; d < 3; ++d) q[1 + d] = nv[d] * a * phi;
        for (int d = 0; d < 3; ++d) q[4 + d] = 1.0 * a * psi;
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) q[7 + 3 * d + e] = -(nv[d] * a * phi);
    }
    double pre1 = 0.5 * (1.0 + prm->eps), pre2 = 0.5 * (1.0 + 1.0 / prm->eps);
    const int32_t ntarg = nt < 0 ? n : nt;
    for (int32_t it = 0; it < ntarg; ++it) {
        const int32_t i = nt < 0 ? it : targets[it];
        c.tx = xyz[3 * i]; c.ty = xyz[3 * i + 1]; c.tz = xyz[3 * i + 2];
        const double *nv = nrm + 3 * (int64_t)i;
        c.tq[0] = 1.0;
        for (int d = 0; d < 3; ++d) c.tq[1 + d] = 1.0;
        for (int d = 0; d < 3; ++d) c.tq[4 + d] = nv[d];
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) c.tq[7 + 3 * d + e] = nv[e];
        c.self = i;
        double peng[2];
        compp_tree(&c, 0, peng);
        if (nt < 0) {
            y[i] = pre1 * x[i] - peng[0];
            y[n + i] = pre2 * x[n + i] - peng[1];
        } else {
            y[2 * it] = pre1 * x[i] - peng[0];
            y[2 * it + 1] = pre2 * x[n + i] - peng[1];
        }
    }
    if (counts) { counts[0] = c.npair_near; counts[1] = c.npair_far; counts[2] = c.nmom; }
    for (int32_t i = 0; i < nn; ++i) free(c.ms[i]);
    free(c.ms); free(c.sq); free(c.dercof);
    tco_free(&c.T);
}

void or_direct_matvec(const or_params *prm, int32_t n, const double *xyz,
                      const double *nrm, const double *area, const double *x, double *y)
{
    double pre1 = 0.5 * (1.0 + prm->eps), pre2 = 0.5 * (1.0 + 1.0 / prm->eps);
    for (int32_t i = 0; i < n; ++i) {
        double peng[2] = {0.0, 0.0};
        for (int32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            pair_pb(prm->kappa, prm->eps, xyz + 3 * (int64_t)i, nrm + 3 * (int64_t)i,
                    xyz + 3 * (int64_t)j, nrm + 3 * (int64_t)j, area[j], x[j], x[n + j], peng);
        }
        y[i] = pre1 * x[i] - peng[0];
        y[n + i] = pre2 * x[n + i] - peng[1];
    }
}

/* ------------------------------------------------------------------------- */
/* leaf-block preconditioner                                                  */
/* ------------------------------------------------------------------------- */
void or_leaf_block(const or_params *prm, const double *xyz, const double *nrm,
                   const double *area, int32_t ibeg, int32_t iend, double *blk)
{
    /* src/treecode3d_pb.f:1439-1528; column-major, leading dimension 2m */
    int32_t m = iend - ibeg + 1, m2 = 2 * m;
    double eps = prm->eps, kappa = prm->kappa;
    double pre1 = 0.5 * (1.0 + eps), pre2 = 0.5 * (1.0 + 1.0 / eps);
    memset(blk, 0, sizeof(double) * (size_t)m2 * m2);
    for (int32_t i = ibeg; i <= iend; ++i) {
        const double *s = xyz + 3 * (int64_t)i, *v0 = nrm + 3 * (int64_t)i;
        for (int32_t j = ibeg; j <= iend; ++j) {
            int64_t ii = i - ibeg, jj = j - ibeg;
            if (j == i) {
                blk[ii * (m2 + 1)] = pre1;
                blk[(ii + m) * (m2 + 1)] = pre2;
                continue;
            }
            const double *r = xyz + 3 * (int64_t)j, *v = nrm + 3 * (int64_t)j;
            double d[3] = {r[0] - s[0], r[1] - s[1], r[2] - s[2]};
            double rs = sqrt(dot3(d, d));
            double g0 = ONE_OVER_4PI / rs;
            double krs = kappa * rs;
            double ek = exp(-krs);
            double gk = ek * g0;
            double ct = dot3(v, d) / rs;
            double ct0 = dot3(v0, d) / rs;
            double tp1 = g0 / rs;
            double tp2 = (1.0 + krs) * ek;
            double g10 = ct0 * tp1, g20 = tp2 * g10;
            double g1 = ct * tp1, g2 = tp2 * g1;
            double g3 = (dot3(v0, v) - 3.0 * ct0 * ct) / rs * tp1;
            double g4 = tp2 * g3 - kappa * kappa * ct0 * ct * gk;
            blk[jj * m2 + ii] = -(g1 - eps * g2) * area[j];
            blk[jj * m2 + ii + m] = -(g0 - gk) * area[j];
            blk[(jj + m) * m2 + ii] = -(g4 - g3) * area[j];
            blk[(jj + m) * m2 + ii + m] = -(g10 - g20 / eps) * area[j];
        }
    }
}

/* partial-pivot LU of a column-major m x m matrix, first-maximum pivot rule
 * (DGETRF/DGETRF2/IDAMAX semantics, src/dgetrf_dep.f:107-222,757-914);
 * returns 0 or 1+column of the first zero pivot */
static int32_t lu_factor(int32_t m, double *a, int32_t *piv)
{
    int32_t info = 0;
    for (int32_t k = 0; k < m; ++k) {
        int32_t p = k;
        double best = fabs(a[(int64_t)k * m + k]);
        for (int32_t i = k + 1; i < m; ++i) {
            double v = fabs(a[(int64_t)k * m + i]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (a[(int64_t)k * m + p] == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        if (p != k)
            for (int32_t j = 0; j < m; ++j) {
                double t = a[(int64_t)j * m + k];
                a[(int64_t)j * m + k] = a[(int64_t)j * m + p];
                a[(int64_t)j * m + p] = t;
            }
        double r = 1.0 / a[(int64_t)k * m + k];
        for (int32_t i = k + 1; i < m; ++i) a[(int64_t)k * m + i] *= r;
        for (int32_t j = k + 1; j < m; ++j) {
            double u = a[(int64_t)j * m + k];
            if (u == 0.0) continue;
            double *cj = a + (int64_t)j * m;
            const double *ck = a + (int64_t)k * m;
            for (int32_t i = k + 1; i < m; ++i) cj[i] = cj[i] - ck[i] * u;
        }
    }
    return info;
}

static void lu_solve(int32_t m, const double *a, const int32_t *piv, double *b)
{
    for (int32_t k = 0; k < m; ++k)
        if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int32_t k = 0; k < m; ++k) {           /* unit lower */
        double bk = b[k];
        const double *ck = a + (int64_t)k * m;
        for (int32_t i = k + 1; i < m; ++i) b[i] = b[i] - bk * ck[i];
    }
    for (int32_t k = m - 1; k >= 0; --k) {      /* upper */
        const double *ck = a + (int64_t)k * m;
        b[k] = b[k] / ck[k];
        double bk = b[k];
        for (int32_t i = 0; i < k; ++i) b[i] = b[i] - bk * ck[i];
    }
}

int32_t or_precond(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                   const double *area, int32_t nleaf, const int32_t *ibeg,
                   const int32_t *iend, const double *r, double *z)
{
    int32_t bad = 0;
    for (int32_t l = 0; l < nleaf; ++l) {
        int32_t m = iend[l] - ibeg[l] + 1, m2 = 2 * m;
        double *blk = malloc(sizeof(double) * (size_t)m2 * m2);
        double *rhs = malloc(sizeof(double) * m2);
        int32_t *piv = malloc(sizeof(int32_t) * m2);
        or_leaf_block(prm, xyz, nrm, area, ibeg[l], iend[l], blk);
        for (int32_t i = 0; i < m; ++i) {
            rhs[i] = r[ibeg[l] + i];
            rhs[m + i] = r[n + ibeg[l] + i];
        }
        if (lu_factor(m2, blk, piv) && !bad) bad = l + 1;
        lu_solve(m2, blk, piv, rhs);
        for (int32_t i = 0; i < m; ++i) {
            z[ibeg[l] + i] = rhs[i];
            z[n + ibeg[l] + i] = rhs[m + i];
        }
        free(blk); free(rhs); free(piv);
    }
    return bad;
}

/* ------------------------------------------------------------------------- */
/* GMRES (SLATEC DGMRES, ITOL=2, JPRE=-1, JSCAL=0, KMP=MAXL)                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    const or_params *prm;
    int32_t n;
    const double *xyz, *nrm, *area;
    const or_node *nodes;
    int32_t nleaf;
    int32_t *lb, *le;
} sys_ctx;

static void sys_matvec(sys_ctx *s, const double *x, double *y)
{
    int64_t n2 = 2 * (int64_t)s->n;
    double sa = 0.0;
    for (int64_t i = 0; i < n2; ++i) sa = sa + fabs(x[i]);
    if (sa < 1.0e-10) {                         /* src/main.f90:478-479 */
        memcpy(y, x, sizeof(double) * n2);
        return;
    }
    or_treecode_matvec(s->prm, s->n, s->xyz, s->nrm, s->area, s->nodes, x, y, NULL);
}

static void sys_msolve(sys_ctx *s, const double *r, double *z)
{
    or_precond(s->prm, s->n, s->xyz, s->nrm, s->area, s->nleaf, s->lb, s->le, r, z);
}

static double vnrm2(int64_t n, const double *v)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = s + v[i] * v[i];
    return sqrt(s);
}

static double vdot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = s + a[i] * b[i];
    return s;
}

/* DORTH, src/dgmres_dep.f:1376-1411 (KMP = MAXL: full MGS from column 1) */
static void dorth(int64_t n, double *vnew, double **v, double *hcol, int ll, double *snormw)
{
    double vnrm = vnrm2(n, vnew);
    for (int i = 0; i < ll; ++i) {
        hcol[i] = vdot(n, v[i], vnew);
        double tem = -hcol[i];
        for (int64_t k = 0; k < n; ++k) vnew[k] = vnew[k] + tem * v[i][k];
    }
    *snormw = vnrm2(n, vnew);
    if (vnrm + 0.001 * *snormw != vnrm) return;
    double sumdsq = 0.0;
    for (int i = 0; i < ll; ++i) {
        double tem = -vdot(n, v[i], vnew);
        if (hcol[i] + 0.001 * tem == hcol[i]) continue;
        hcol[i] = hcol[i] - tem;
        for (int64_t k = 0; k < n; ++k) vnew[k] = vnew[k] + tem * v[i][k];
        sumdsq = sumdsq + tem * tem;
    }
    if (sumdsq == 0.0) return;
    double arg = *snormw * *snormw - sumdsq;
    if (arg < 0.0) arg = 0.0;
    *snormw = sqrt(arg);
}

/* DHEQR update step (IJOB > 1), src/dgmres_dep.f:1093-1125; hes column-major
 * with leading dimension ldh; columns/rows 0-based; returns info */
static int dheqr_update(double *hes, int ldh, int nn, double *q)
{
    /* nn = current column count (1-based N in the reference) */
    double *col = hes + (int64_t)(nn - 1) * ldh;
    for (int k = 0; k < nn - 1; ++k) {
        double c = q[2 * k], s = q[2 * k + 1];
        double t1 = col[k], t2 = col[k + 1];
        col[k] = c * t1 - s * t2;
        col[k + 1] = s * t1 + c * t2;
    }
    double t1 = col[nn - 1], t2 = col[nn], c, s;
    if (t2 == 0.0) { c = 1.0; s = 0.0; }
    else if (fabs(t2) >= fabs(t1)) {
        double t = t1 / t2;
        s = -1.0 / sqrt(1.0 + t * t);
        c = -s * t;
    } else {
        double t = t2 / t1;
        c = 1.0 / sqrt(1.0 + t * t);
        s = -c * t;
    }
    q[2 * (nn - 1)] = c;
    q[2 * (nn - 1) + 1] = s;
    col[nn - 1] = c * t1 - s * t2;
    return col[nn - 1] == 0.0 ? nn : 0;
}

/* DHELS, src/dgmres_dep.f:925-947 */
static void dhels(const double *hes, int ldh, int nn, const double *q, double *b)
{
    for (int k = 0; k < nn; ++k) {
        double c = q[2 * k], s = q[2 * k + 1];
        double t1 = b[k], t2 = b[k + 1];
        b[k] = c * t1 - s * t2;
        b[k + 1] = s * t1 + c * t2;
    }
    for (int kb = 0; kb < nn; ++kb) {
        int k = nn - 1 - kb;
        b[k] = b[k] / hes[(int64_t)k * ldh + k];
        double t = -b[k];
        for (int i = 0; i < k; ++i) b[i] = b[i] + t * hes[(int64_t)k * ldh + i];
    }
}

/* DRLCAL with KMP = MAXL, src/dgmres_dep.f:1948-1977 */
static void drlcal(int64_t n, int ll, double **v, const double *q, double *rl,
                   double snormw, double prod, double r0nrm)
{
    memcpy(rl, v[0], sizeof(double) * n);
    for (int i = 1; i <= ll - 1; ++i) {
        double s = q[2 * i - 1], c = q[2 * i - 2];
        for (int64_t k = 0; k < n; ++k) rl[k] = s * rl[k] + c * v[i][k];
    }
    double s = q[2 * ll - 1], c = q[2 * ll - 2] / snormw;
    for (int64_t k = 0; k < n; ++k) rl[k] = s * rl[k] + c * v[ll][k];
    double tem = r0nrm * prod;
    for (int64_t k = 0; k < n; ++k) rl[k] = tem * rl[k];
}

int32_t or_gmres(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                 const double *area, const or_node *nodes, const double *b, double *x,
                 double tol, int32_t maxl, int32_t nrmax, int32_t *iter_out, double *err_out,
                 double *err_hist)
{
    sys_ctx S;
    S.prm = prm; S.n = n; S.xyz = xyz; S.nrm = nrm; S.area = area; S.nodes = nodes;
    S.nleaf = or_leaves(nodes, NULL, NULL);
    S.lb = malloc(sizeof(int32_t) * S.nleaf);
    S.le = malloc(sizeof(int32_t) * S.nleaf);
    or_leaves(nodes, S.lb, S.le);

    int64_t N = 2 * (int64_t)n;
    if (maxl > N) maxl = (int32_t)N;
    int ldh = maxl + 1;
    double **v = malloc(sizeof(double *) * (maxl + 1));
    for (int i = 0; i <= maxl; ++i) v[i] = malloc(sizeof(double) * N);
    double *r = malloc(sizeof(double) * N), *z = malloc(sizeof(double) * N);
    double *wk = malloc(sizeof(double) * N), *dl = malloc(sizeof(double) * N);
    double *hes = malloc(sizeof(double) * ldh * maxl);
    double *q = malloc(sizeof(double) * 2 * maxl);
    double *yls = malloc(sizeof(double) * (maxl + 1));
    int32_t iter = 0, ierr = 0, nhist = 0;
    double err = 0.0;

    /* DGMRES, src/dgmres_dep.f:754-815 */
    sys_msolve(&S, b, r);
    double bnrm = vnrm2(N, r);
    sys_matvec(&S, x, r);
    for (int64_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
    int nrsts = 0;
    for (;;) {
        if (nrsts > nrmax) { ierr = 1; break; }
        if (nrsts > 0) memcpy(r, dl, sizeof(double) * N);
        /* DPIGMR, src/dgmres_dep.f:1652-1863 */
        memset(z, 0, sizeof(double) * N);
        int iflag = 0, lgmr = 0;
        if (nrsts == 0) {
            memcpy(wk, r, sizeof(double) * N);
            sys_msolve(&S, wk, r);
        }
        memcpy(v[0], r, sizeof(double) * N);
        double r0nrm = vnrm2(N, v[0]);
        err = vnrm2(N, v[0]) / bnrm;           /* ISDGMR with LGMR = 0 */
        if (err <= tol) {
            iflag = 0;
            goto pigmr_done;
        }
        {
            double tem = 1.0 / r0nrm;
            for (int64_t k = 0; k < N; ++k) v[0][k] = tem * v[0][k];
            memset(hes, 0, sizeof(double) * ldh * maxl);
            double prod = 1.0, rho = 0.0, snormw = 0.0;
            int converged = 0, ll;
            for (ll = 1; ll <= maxl; ++ll) {
                lgmr = ll;
                memcpy(wk, v[ll - 1], sizeof(double) * N);
                sys_matvec(&S, wk, v[ll]);
                memcpy(wk, v[ll], sizeof(double) * N);
                sys_msolve(&S, wk, v[ll]);
                double *hcol = hes + (int64_t)(ll - 1) * ldh;
                dorth(N, v[ll], v, hcol, ll, &snormw);
                hcol[ll] = snormw;
                if (dheqr_update(hes, ldh, ll, q) == ll) { iflag = 2; goto stalled; }
                prod = prod * q[2 * ll - 1];
                rho = fabs(prod * r0nrm);
                /* ISDGMR: residual via DRLCAL, err = ||rl|| / bnrm */
                drlcal(N, ll, v, q, dl, snormw, prod, r0nrm);
                err = vnrm2(N, dl) / bnrm;
                if (err_hist) err_hist[nhist++] = err;
                if (err <= tol) { converged = 1; break; }
                if (ll == maxl) break;
                tem = 1.0 / snormw;
                for (int64_t k = 0; k < N; ++k) v[ll][k] = tem * v[ll][k];
            }
            if (!converged) {
                if (rho < r0nrm) {
                    iflag = 1;
                    drlcal(N, maxl, v, q, dl, snormw, prod, r0nrm);
                } else {
                    iflag = 2;
                    goto stalled;
                }
            }
            /* label 200: form the update z = V y */
            int nn = lgmr;
            for (int k = 0; k <= nn; ++k) yls[k] = 0.0;
            yls[0] = r0nrm;
            dhels(hes, ldh, nn, q, yls);
            for (int i = 0; i < nn; ++i)
                for (int64_t k = 0; k < N; ++k) z[k] = z[k] + yls[i] * v[i][k];
            goto pigmr_done;
        }
    stalled:
        memset(z, 0, sizeof(double) * N);
    pigmr_done:
        iter += lgmr;
        for (int64_t k = 0; k < N; ++k) x[k] = x[k] + z[k];
        if (iflag == 0) { ierr = 0; break; }
        if (iflag == 1) { ++nrsts; continue; }
        ierr = 2;
        break;
    }
    *iter_out = iter;
    *err_out = err;
    for (int i = 0; i <= maxl; ++i) free(v[i]);
    free(v); free(r); free(z); free(wk); free(dl); free(hes); free(q); free(yls);
    free(S.lb); free(S.le);
    return ierr;
}

/* ------------------------------------------------------------------------- */
/* energy                                                                     */
/* ------------------------------------------------------------------------- */
double or_energy(int32_t n, const double *xyz, const double *nrm, const double *area,
                 const double *x, int32_t nq, const double *qpos, const double *q,
                 double eps, double kappa, double eps_p, double *couleng)
{
    const double para = 332.0716;                 /* src/main.f90:64 */
    double sol = 0.0, cou = 0.0;
    for (int32_t k = 0; k < nq; ++k) {
        const double *r0 = qpos + 3 * (int64_t)k;
        double ptl = 0.0;
        for (int32_t j = 0; j < n; ++j) {         /* potential_molecule, main.f90:627-655 */
            const double *r = xyz + 3 * (int64_t)j, *v = nrm + 3 * (int64_t)j;
            double d[3] = {r[0] - r0[0], r[1] - r0[1], r[2] - r0[2]};
            double rs = sqrt(dot3(d, d));
            double g0 = ONE_OVER_4PI / rs;
            double krs = kappa * rs;
            double ek = exp(-krs);
            double gk = ek * g0;
            double ct = dot3(v, d) / rs;
            double tp1 = g0 / rs;
            double tp2 = (1.0 + krs) * ek;
            double g1 = ct * tp1, g2 = tp2 * g1;
            double h1 = g1 - eps * g2, h2 = g0 - gk;
            ptl = ptl + area[j] * h1 * x[j];
            ptl = ptl + area[j] * h2 * x[n + j];
        }
        sol = sol + q[k] * ptl;
        for (int32_t j = k + 1; j < nq; ++j) {
            const double *r1 = qpos + 3 * (int64_t)j;
            double d[3] = {r0[0] - r1[0], r0[1] - r1[1], r0[2] - r1[2]};
            double dist = sqrt(dot3(d, d));
            cou = cou + 1.0 / eps_p / dist * q[k] * q[j];
        }
    }
    sol = sol * 0.5 * para * 4.0 * PI;
    if (couleng) *couleng = cou * para;
    return sol;
}

/* ------------------------------------------------------------------------- */
/* post-processing and the other preconditioner                               */
/* ------------------------------------------------------------------------- */
void or_direct_rows(const or_params *prm, int32_t n, const double *xyz, const double *nrm,
                    const double *area, const double *x, int32_t nt, const int32_t *targets,
                    double *y)
{
    /* rows of the O(n^2) direct operator (the treecode at theta = 0,
     * src/treecode3d_pb.f:1158-1185 with every source direct) for the listed
     * targets; y[2k], y[2k+1] for target k.  Targets are independent, so they
     * are spread over threads (each row's sum keeps the serial order). */
    double pre1 = 0.5 * (1.0 + prm->eps), pre2 = 0.5 * (1.0 + 1.0 / prm->eps);
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t k = 0; k < nt; ++k) {
        const int32_t i = targets[k];
        double peng[2] = {0.0, 0.0};
        for (int32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            pair_pb(prm->kappa, prm->eps, xyz + 3 * (int64_t)i, nrm + 3 * (int64_t)i,
                    xyz + 3 * (int64_t)j, nrm + 3 * (int64_t)j, area[j], x[j], x[n + j], peng);
        }
        y[2 * (int64_t)k] = pre1 * x[i] - peng[0];
        y[2 * (int64_t)k + 1] = pre2 * x[n + i] - peng[1];
    }
}

void or_diag_prec(int32_t n2, double eps, const double *r, double *z)
{
    /* MSOLVEprec, src/main.f90:532-542 */
    double scale1 = 0.5 * (1.0 + eps), scale2 = 0.5 * (1.0 + 1.0 / eps);
    for (int32_t i = 0; i < n2 / 2; ++i) z[i] = r[i] / scale1;
    for (int32_t i = n2 / 2; i < n2; ++i) z[i] = r[i] / scale2;
}

int32_t or_vertex_potential(int32_t n, int32_t nvert, const double *vpos, const int32_t *faces1,
                            const double *xyz, const double *x, double *vpot)
{
    /* output_potential, src/main.f90:226-326, in the original face order
     * (xyz: centroids, x: solution un-permuted as at main.f90:249-254).  Each
     * vertex keeps the first 14 faces that list it (ind_vert, main.f90:257-268),
     * then the 1/distance-weighted means (main.f90:270-288), scaled by
     * para*4*pi (main.f90:245,289).  Returns 1 for a vertex index out of range. */
    enum { NSLOT = 14 };
    const double para_temp = 332.0716 * 4 * PI;
    int32_t *cnt = calloc((size_t)nvert, sizeof(int32_t));
    int32_t *slot = malloc(sizeof(int32_t) * (size_t)nvert * NSLOT);
    if (!cnt || !slot) { free(cnt); free(slot); return 2; }
    for (int32_t i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            int32_t v = faces1[3 * (int64_t)i + j] - 1;
            if (v < 0 || v >= nvert) { free(cnt); free(slot); return 1; }
            if (cnt[v] < NSLOT) slot[(int64_t)v * NSLOT + cnt[v]++] = i;
        }
    for (int32_t v = 0; v < nvert; ++v) {
        double s1 = 0.0, s2 = 0.0, tot = 0.0;
        const double *p = vpos + 3 * (int64_t)v;
        for (int32_t k = 0; k < cnt[v]; ++k) {
            int32_t f = slot[(int64_t)v * NSLOT + k];
            const double *c = xyz + 3 * (int64_t)f;
            double aa[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
            double len = sqrt(dot3(aa, aa));
            s1 = s1 + 1.0 / len * x[f];
            s2 = s2 + 1.0 / len * x[f + n];
            tot = tot + 1.0 / len;
        }
        vpot[2 * (int64_t)v] = s1 / tot * para_temp;
        vpot[2 * (int64_t)v + 1] = s2 / tot * para_temp;
    }
    free(cnt);
    free(slot);
    return 0;
}
