// Host-side tree build over face centroids.
//
// Restates the reference's recursive box bisection (CREATE_TREE/PARTITION_8,
// src/treecode3d_pb.f:142-356) and its in-place hole-walk partition (PARTITION,
// src/treecode3d_pb.f:1677-1759).  The integer permutation it produces must be
// bit-identical to the reference's orderarr, so the partition keeps the
// reference's exact element movement (set the first element aside, alternate
// scans from the top for an element <= the split value and from the bottom for
// one > it, moving each into the travelling hole) and every floating-point
// expression that feeds a comparison (box midpoints, extents, critical length)
// is evaluated in the reference's operand order without contraction.
#include <cmath>
#include <cstring>

#include "tabi_internal.h"

#pragma clang fp contract(off)

namespace tabi {
namespace {

struct Ctx {
    double *x, *y, *z;
    int32_t *ord;
    int32_t maxpar;
    HostTree *t;
    std::vector<std::vector<int32_t>> kids;
};

// Returns midind with a[ibeg..midind] <= val < a[midind+1..iend].
int32_t partition(double *a, double *b, double *c, int32_t *ind, int32_t ibeg,
                  int32_t iend, double val)
{
    if (ibeg < iend) {
        const double ta = a[ibeg], tb = b[ibeg], tc = c[ibeg];
        const int32_t tind = ind[ibeg];
        a[ibeg] = val;  // sentinel: the scan from the bottom stops on it
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
        const int32_t midind = (ta > val) ? upper - 1 : upper;
        a[upper] = ta; b[upper] = tb; c[upper] = tc; ind[upper] = tind;
        return midind;
    }
    if (ibeg == iend) return a[ibeg] <= val ? ibeg : ibeg - 1;
    return ibeg - 1;
}

void range_minmax(const double *a, int32_t i0, int32_t i1, double &mn, double &mx)
{
    double lo = a[i0], hi = a[i0];
    for (int32_t i = i0 + 1; i <= i1; ++i) {
        if (a[i] < lo) lo = a[i];
        if (a[i] > hi) hi = a[i];
    }
    mn = lo;
    mx = hi;
}

int32_t create(Ctx &c, int32_t ibeg, int32_t iend, const double bx[6], int32_t level)
{
    HostTree &t = *c.t;
    const int32_t id = t.nnode();
    const double xl = bx[1] - bx[0], yl = bx[3] - bx[2], zl = bx[5] - bx[4];
    double lmax = xl;
    if (yl > lmax) lmax = yl;
    if (zl > lmax) lmax = zl;
    const double xm = (bx[1] + bx[0]) / 2.0, ym = (bx[3] + bx[2]) / 2.0,
                 zm = (bx[5] + bx[4]) / 2.0;
    const double t1 = bx[1] - xm, t2 = bx[3] - ym, t3 = bx[5] - zm;
    t.ibeg.push_back(ibeg); t.iend.push_back(iend); t.numpar.push_back(iend - ibeg + 1);
    t.level.push_back(level); t.nchild.push_back(0); t.child_off.push_back(0);
    t.xmin.push_back(bx[0]); t.xmax.push_back(bx[1]); t.ymin.push_back(bx[2]);
    t.ymax.push_back(bx[3]); t.zmin.push_back(bx[4]); t.zmax.push_back(bx[5]);
    t.xmid.push_back(xm); t.ymid.push_back(ym); t.zmid.push_back(zm);
    t.radius.push_back(std::sqrt(t1 * t1 + t2 * t2 + t3 * t3));
    c.kids.emplace_back();
    if (iend - ibeg + 1 <= c.maxpar) {
        t.leaves.push_back(id);
        return id;
    }
    int32_t ind[8][2];
    double bxs[8][6];
    std::memset(ind, 0, sizeof(ind));
    std::memcpy(bxs[0], bx, sizeof(bxs[0]));
    ind[0][0] = ibeg;
    ind[0][1] = iend;
    const double critlen = lmax / std::sqrt(2.0);
    int nch = 1;
    if (xl >= critlen) {
        const int32_t mid = partition(c.x, c.y, c.z, c.ord, ind[0][0], ind[0][1], xm);
        ind[1][0] = mid + 1; ind[1][1] = ind[0][1]; ind[0][1] = mid;
        std::memcpy(bxs[1], bxs[0], sizeof(bxs[0]));
        bxs[0][1] = xm; bxs[1][0] = xm;
        nch *= 2;
    }
    if (yl >= critlen) {
        for (int i = 0; i < nch; ++i) {
            const int32_t mid = partition(c.y, c.x, c.z, c.ord, ind[i][0], ind[i][1], ym);
            ind[nch + i][0] = mid + 1; ind[nch + i][1] = ind[i][1]; ind[i][1] = mid;
            std::memcpy(bxs[nch + i], bxs[i], sizeof(bxs[0]));
            bxs[i][3] = ym; bxs[nch + i][2] = ym;
        }
        nch *= 2;
    }
    if (zl >= critlen) {
        for (int i = 0; i < nch; ++i) {
            const int32_t mid = partition(c.z, c.x, c.y, c.ord, ind[i][0], ind[i][1], zm);
            ind[nch + i][0] = mid + 1; ind[nch + i][1] = ind[i][1]; ind[i][1] = mid;
            std::memcpy(bxs[nch + i], bxs[i], sizeof(bxs[0]));
            bxs[i][5] = zm; bxs[nch + i][4] = zm;
        }
        nch *= 2;
    }
    // children holding two or more particles get the tight box of their
    // particles; single-particle children keep the split box
    for (int i = 0; i < 8; ++i) {
        if (ind[i][0] < ind[i][1]) {
            range_minmax(c.x, ind[i][0], ind[i][1], bxs[i][0], bxs[i][1]);
            range_minmax(c.y, ind[i][0], ind[i][1], bxs[i][2], bxs[i][3]);
            range_minmax(c.z, ind[i][0], ind[i][1], bxs[i][4], bxs[i][5]);
        }
    }
    for (int i = 0; i < nch; ++i) {
        if (ind[i][0] <= ind[i][1]) {
            const int32_t ch = create(c, ind[i][0], ind[i][1], bxs[i], level + 1);
            c.kids[id].push_back(ch);
        }
    }
    return id;
}

}  // namespace

void build_tree_host(int32_t n, double *x, double *y, double *z, int32_t maxparnode,
                     HostTree &t)
{
    t = HostTree();
    t.n = n;
    t.order.resize(n);
    for (int32_t i = 0; i < n; ++i) t.order[i] = i;
    Ctx c{x, y, z, t.order.data(), maxparnode, &t, {}};
    double bx[6];
    range_minmax(x, 0, n - 1, bx[0], bx[1]);
    range_minmax(y, 0, n - 1, bx[2], bx[3]);
    range_minmax(z, 0, n - 1, bx[4], bx[5]);
    create(c, 0, n - 1, bx, 0);
    const int32_t nn = t.nnode();
    for (int32_t i = 0; i < nn; ++i) {
        t.child_off[i] = (int32_t)t.children.size();
        t.nchild[i] = (int32_t)c.kids[i].size();
        for (int32_t k : c.kids[i]) t.children.push_back(k);
    }
}

}  // namespace tabi
