// GPU tree build over face centroids, bit-identical to the reference's.
//
// The reference builds the tree recursively (CREATE_TREE / PARTITION_8,
// src/treecode3d_pb.f:142-356) and permutes the particles with a sequential
// hole-walk partition (PARTITION, src/treecode3d_pb.f:1677-1759).  Here the
// nodes of one tree level are split together.  Each partition pass moves every
// element to the place the hole walk would put it, computed from prefix counts
// (the closed form is checked against the literal hole walk in
// tests/test_partition_form.py):
//
//   segment a[b..e] split at val, first element set aside, "small" = (<= val):
//   c = #small in a[b+1..e], H = b + c;  smalls at positions >= H ranked from
//   the right s_1 > s_2 > ..., bigs at positions b+1..H ranked from the left
//   b_1 < b_2 < ...;  s_1 -> b, s_(j+1) -> b_j, b_j -> s_j, a[b] -> H;
//   midind = H if a[b] <= val else H - 1.
//
// Per pass: k_part_flags (small flags) -> three-kernel inclusive scan ->
// k_part_classify (positions of ranked smalls/bigs, H and midind per segment)
// -> k_part_move (scatter of x, y, z and the index into the other buffer).
// The per-node logic (box, critical length, which halves to split, child
// creation order) runs on the host exactly as the reference orders it; boxes
// of the children come from k_seg_minmax.  Comparisons and the values compared
// against (box midpoints) are exact, so the permutation is the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tabi_internal.h"

#pragma clang fp contract(off)

namespace tabi {
namespace {

#define TCHK(expr)                                                                    \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Seg {
    int32_t b, e;     // inclusive range; e >= b
    double val;       // split value
};

constexpr int kMaxDepth = 200;   // the reference recursion has no bound; coincident
                                 // centroids beyond maxparnode would never split
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanBlock = kScanThreads * kScanPer;

__device__ __forceinline__ int find_seg(const Seg *segs, int nseg, int p)
{
    // segments are sorted by b and disjoint; largest s with segs[s].b <= p
    int lo = 0, hi = nseg - 1, s = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].b <= p) { s = mid; lo = mid + 1; } else hi = mid - 1;
    }
    if (s >= 0 && p > segs[s].e) s = -1;
    return s;
}

__global__ void k_part_flags(int n, const double *__restrict__ a, const Seg *__restrict__ segs,
                             int nseg, int32_t *__restrict__ flag)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = find_seg(segs, nseg, p);
    flag[p] = (s >= 0 && p > segs[s].b && a[p] <= segs[s].val) ? 1 : 0;
}

// inclusive scan, three passes
__global__ void __launch_bounds__(kScanThreads) k_scan_blocks(int n, const int32_t *__restrict__ in,
                                                              int32_t *__restrict__ out,
                                                              int32_t *__restrict__ bsum)
{
    __shared__ int32_t sh[kScanThreads];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    int32_t v[kScanPer];
    int32_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        run += (base + k < n) ? in[base + k] : 0;
        v[k] = run;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int32_t excl = threadIdx.x ? sh[threadIdx.x - 1] : 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < n) out[base + k] = v[k] + excl;
    if (threadIdx.x == kScanThreads - 1) bsum[blockIdx.x] = sh[threadIdx.x];
}

__global__ void __launch_bounds__(1024) k_scan_bsums(int nb, int32_t *__restrict__ bsum)
{
    __shared__ int32_t sh[1024];
    const int t = threadIdx.x;
    int32_t carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        sh[t] = (c0 + t < nb) ? bsum[c0 + t] : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int32_t v = t >= off ? sh[t - off] : 0;
            __syncthreads();
            sh[t] += v;
            __syncthreads();
        }
        if (c0 + t < nb) bsum[c0 + t] = sh[t] + carry;  // inclusive
        const int32_t tot = sh[1023];
        __syncthreads();
        carry += tot;
    }
}

__global__ void k_scan_add(int n, const int32_t *__restrict__ bsum, int32_t *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int blk = p / kScanBlock;
    if (blk > 0) out[p] += bsum[blk - 1];
}

__global__ void k_part_classify(int n, const double *__restrict__ a, const Seg *__restrict__ segs,
                                int nseg, const int32_t *__restrict__ G,
                                const int32_t *__restrict__ flag, int32_t *__restrict__ spos,
                                int32_t *__restrict__ bpos, int32_t *__restrict__ segout)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = find_seg(segs, nseg, p);
    if (s < 0) return;
    const int b = segs[s].b, e = segs[s].e;
    const int32_t g0 = G[b];
    const int c = G[e] - g0;
    const int H = b + c;
    if (p == b) {
        segout[2 * s] = H;
        segout[2 * s + 1] = (a[b] <= segs[s].val) ? H : H - 1;
        return;
    }
    const int S = G[p] - g0;
    if (flag[p]) {
        if (p >= H) spos[b + (c - S + 1)] = p;
    } else if (p <= H) {
        bpos[b + ((p - b) - S)] = p;
    }
}

__global__ void k_part_move(int n, const Seg *__restrict__ segs, int nseg,
                            const int32_t *__restrict__ G, const int32_t *__restrict__ flag,
                            const int32_t *__restrict__ spos, const int32_t *__restrict__ bpos,
                            const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const int32_t *__restrict__ ord,
                            double *__restrict__ x2, double *__restrict__ y2,
                            double *__restrict__ z2, int32_t *__restrict__ ord2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int dest = p;
    const int s = find_seg(segs, nseg, p);
    if (s >= 0) {
        const int b = segs[s].b, e = segs[s].e;
        const int32_t g0 = G[b];
        const int c = G[e] - g0;
        const int H = b + c;
        if (p == b) {
            dest = H;
        } else {
            const int S = G[p] - g0;
            if (flag[p]) {
                if (p >= H) {
                    const int r = c - S + 1;
                    dest = (r == 1) ? b : bpos[b + r - 1];
                }
            } else if (p <= H) {
                dest = spos[b + ((p - b) - S)];
            }
        }
    }
    x2[dest] = x[p];
    y2[dest] = y[p];
    z2[dest] = z[p];
    ord2[dest] = ord[p];
}

// min/max of x, y, z over each range [b, e], in chunks of kMinMaxChunk points:
// workgroup (r, c) reduces chunk c of range r into part[(r*gridDim.y + c)*6 ..]
// (chunks past the range's end write nothing); k_seg_minmax_fin combines them.
// Min and max do not depend on the order, so the boxes are exact.
constexpr int kMinMaxChunk = 4096;
__global__ void __launch_bounds__(256) k_seg_minmax(const int32_t *__restrict__ rng,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    double *__restrict__ part)
{
    __shared__ double sh[6][256];
    const int r = blockIdx.x, c = blockIdx.y, t = threadIdx.x;
    const int b0 = rng[2 * r], e0 = rng[2 * r + 1];
    const int b = b0 + c * kMinMaxChunk, e = min(e0, b + kMinMaxChunk - 1);
    if (c > 0 && b > e0) return;   // chunk 0 always writes (+-inf for an empty range)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = b + t; p <= e; p += 256) {
        const double v[3] = {x[p], y[p], z[p]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = v[d] < mn[d] ? v[d] : mn[d];
            mx[d] = v[d] > mx[d] ? v[d] : mx[d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) { sh[2 * d][t] = mn[d]; sh[2 * d + 1][t] = mx[d]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                sh[2 * d][t] = fmin(sh[2 * d][t], sh[2 * d][t + w]);
                sh[2 * d + 1][t] = fmax(sh[2 * d + 1][t], sh[2 * d + 1][t + w]);
            }
        }
        __syncthreads();
    }
    if (t < 6) part[((size_t)r * gridDim.y + c) * 6 + t] = sh[t][0];
}

// out[6r + k] = min (k even) or max (k odd) over range r's chunk partials
__global__ void k_seg_minmax_fin(int nr, int nch, const int32_t *__restrict__ rng,
                                 const double *__restrict__ part, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * nr) return;
    const int r = i / 6, k = i % 6;
    const int len = rng[2 * r + 1] - rng[2 * r];
    const int cnt = len < 0 ? 1 : len / kMinMaxChunk + 1;
    const double *p = part + (size_t)r * nch * 6 + k;
    double v = p[0];
    for (int c = 1; c < cnt; ++c) v = (k & 1) ? fmax(v, p[(size_t)c * 6]) : fmin(v, p[(size_t)c * 6]);
    out[i] = v;
}

// identity permutation
__global__ void k_iota(int n, int32_t *__restrict__ a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Waits for an event by polling (a blocking wait wakes tens of microseconds
// late, once per tree level).
void spin_wait(hipEvent_t ev)
{
    for (;;) {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipSuccess) return;
        if (q != hipErrorNotReady) TCHK(q);
    }
}

template <typename T>
struct Buf {
    T *p = nullptr;
    explicit Buf(size_t n) { p = static_cast<T *>(device_alloc(std::max<size_t>(n, 1) * sizeof(T))); }
    ~Buf() { device_free(p); }
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
};

inline unsigned nb_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct LNode {          // node under construction (level order)
    int32_t ibeg, iend, level;
    double bx[6];
    int parent;
    std::vector<int> kids;
};

}  // namespace

void build_tree_gpu(int32_t n, const double *d_x, const double *d_y, const double *d_z,
                    int32_t maxparnode, void *stream, HostTree &t, double *d_tx,
                    double *d_ty, double *d_tz, int32_t *d_order)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    Buf<double> xb(n), yb(n), zb(n);
    Buf<int32_t> ob(n), ob2(n), flag(n), G(n), spos(n + 2), bpos(n + 2);
    const int nscan = (n + kScanBlock - 1) / kScanBlock;
    Buf<int32_t> bsum(nscan);
    double *X[2] = {d_tx, xb.p}, *Y[2] = {d_ty, yb.p}, *Z[2] = {d_tz, zb.p};
    int32_t *O[2] = {d_order, ob.p};
    int cur = 0;
    TCHK(hipMemcpyAsync(X[0], d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    TCHK(hipMemcpyAsync(Y[0], d_y, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    TCHK(hipMemcpyAsync(Z[0], d_z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_iota<<<nb_of(n, 256), 256, 0, s>>>(n, O[0]);
    TCHK(hipGetLastError());
    const size_t mseg = 8 * (size_t)n / std::max(maxparnode, 1) + 16;   // segments / ranges per level
    Buf<Seg> dseg(mseg);
    Buf<int32_t> dsegout(2 * mseg);
    Buf<int32_t> drng(2 * mseg);
    Buf<double> dmm(6 * mseg);
    std::unique_ptr<Buf<double>> dmmp;     // chunk partials, grown on demand
    size_t dmmp_cap = 0;
    // the per-level tables travel through pinned host memory (inputs in the
    // first half, results in the second); each round trip ends with a polled
    // event, after which the buffer may be rewritten
    const size_t half = std::max(sizeof(Seg), 6 * sizeof(double)) * mseg;
    struct Pinned {
        char *p = nullptr;
        ~Pinned() { if (p) (void)hipHostFree(p); }
    } pin;
    TCHK(hipHostMalloc(reinterpret_cast<void **>(&pin.p), 2 * half, hipHostMallocDefault));
    char *const pin_in = pin.p, *const pin_out = pin.p + half;
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() { if (e) (void)hipEventDestroy(e); }
    } ev;
    TCHK(hipEventCreateWithFlags(&ev.e, hipEventDisableTiming));
    auto round_trip_end = [&]() {
        TCHK(hipEventRecord(ev.e, s));
        spin_wait(ev.e);
    };

    auto minmax = [&](const std::vector<int32_t> &rng, std::vector<double> &out) {
        const int nr = (int)rng.size() / 2;
        out.assign(6 * nr, 0.0);
        if (!nr) return;
        int nch = 1;
        for (int r = 0; r < nr; ++r)
            nch = std::max(nch, std::max(rng[2 * r + 1] - rng[2 * r], 0) / kMinMaxChunk + 1);
        const size_t need = (size_t)6 * nr * nch;
        if (need > dmmp_cap) {   // every earlier use has completed (round trips end with a wait)
            dmmp.reset();
            dmmp.reset(new Buf<double>(need));
            dmmp_cap = need;
        }
        std::memcpy(pin_in, rng.data(), sizeof(int32_t) * rng.size());
        TCHK(hipMemcpyAsync(drng.p, pin_in, sizeof(int32_t) * rng.size(), hipMemcpyHostToDevice, s));
        k_seg_minmax<<<dim3(nr, nch), 256, 0, s>>>(drng.p, X[cur], Y[cur], Z[cur], dmmp->p);
        k_seg_minmax_fin<<<nb_of(6 * nr, 256), 256, 0, s>>>(nr, nch, drng.p, dmmp->p, dmm.p);
        TCHK(hipGetLastError());
        TCHK(hipMemcpyAsync(pin_out, dmm.p, sizeof(double) * 6 * nr, hipMemcpyDeviceToHost, s));
        round_trip_end();
        std::memcpy(out.data(), pin_out, sizeof(double) * 6 * nr);
    };

    // one partition pass over the given segments on coordinate d (0 x, 1 y, 2 z);
    // returns midind per segment (segments with b > e are resolved here)
    auto pass = [&](int d, const std::vector<Seg> &segs_in, std::vector<int32_t> &mid) {
        mid.assign(segs_in.size(), 0);
        std::vector<int> idx;
        std::vector<Seg> segs;
        for (size_t k = 0; k < segs_in.size(); ++k) {
            if (segs_in[k].b > segs_in[k].e) mid[k] = segs_in[k].b - 1;
            else { idx.push_back((int)k); segs.push_back(segs_in[k]); }
        }
        if (segs.empty()) return;
        std::vector<int> perm(segs.size());
        for (size_t k = 0; k < perm.size(); ++k) perm[k] = (int)k;
        std::sort(perm.begin(), perm.end(), [&](int a, int b) { return segs[a].b < segs[b].b; });
        std::vector<Seg> sorted(segs.size());
        for (size_t k = 0; k < perm.size(); ++k) sorted[k] = segs[perm[k]];
        const int ns = (int)sorted.size();
        std::memcpy(pin_in, sorted.data(), sizeof(Seg) * ns);
        TCHK(hipMemcpyAsync(dseg.p, pin_in, sizeof(Seg) * ns, hipMemcpyHostToDevice, s));
        const double *A = d == 0 ? X[cur] : d == 1 ? Y[cur] : Z[cur];
        k_part_flags<<<nb_of(n, 256), 256, 0, s>>>(n, A, dseg.p, ns, flag.p);
        k_scan_blocks<<<nscan, kScanThreads, 0, s>>>(n, flag.p, G.p, bsum.p);
        k_scan_bsums<<<1, 1024, 0, s>>>(nscan, bsum.p);
        k_scan_add<<<nb_of(n, 256), 256, 0, s>>>(n, bsum.p, G.p);
        k_part_classify<<<nb_of(n, 256), 256, 0, s>>>(n, A, dseg.p, ns, G.p, flag.p, spos.p,
                                                      bpos.p, dsegout.p);
        const int nx = cur ^ 1;
        k_part_move<<<nb_of(n, 256), 256, 0, s>>>(n, dseg.p, ns, G.p, flag.p, spos.p, bpos.p,
                                                  X[cur], Y[cur], Z[cur], O[cur], X[nx], Y[nx],
                                                  Z[nx], O[nx]);
        TCHK(hipGetLastError());
        cur = nx;
        TCHK(hipMemcpyAsync(pin_out, dsegout.p, sizeof(int32_t) * 2 * ns, hipMemcpyDeviceToHost, s));
        round_trip_end();
        const int32_t *so = reinterpret_cast<const int32_t *>(pin_out);
        for (int k = 0; k < ns; ++k) mid[idx[perm[k]]] = so[2 * k + 1];
    };

    // root box (SETUP, src/treecode3d_pb.f:120-125)
    std::vector<LNode> nodes;
    {
        std::vector<double> mm;
        minmax({0, n - 1}, mm);
        LNode r;
        r.ibeg = 0; r.iend = n - 1; r.level = 0; r.parent = -1;
        std::memcpy(r.bx, mm.data(), sizeof(r.bx));
        nodes.push_back(r);
    }
    std::vector<int> level_nodes{0};
    while (!level_nodes.empty()) {
        struct Split {
            int node;
            double xm, ym, zm, critlen, xl, yl, zl;
            int32_t ind[8][2];
            double bxs[8][6];
            int nch;
        };
        std::vector<Split> sp;
        for (int id : level_nodes) {
            const LNode &p = nodes[id];
            if (p.iend - p.ibeg + 1 <= maxparnode) continue;
            Split q;
            q.node = id;
            const double *bx = p.bx;
            q.xl = bx[1] - bx[0]; q.yl = bx[3] - bx[2]; q.zl = bx[5] - bx[4];
            double lmax = q.xl;
            if (q.yl > lmax) lmax = q.yl;
            if (q.zl > lmax) lmax = q.zl;
            q.xm = (bx[1] + bx[0]) / 2.0; q.ym = (bx[3] + bx[2]) / 2.0; q.zm = (bx[5] + bx[4]) / 2.0;
            q.critlen = lmax / std::sqrt(2.0);
            std::memset(q.ind, 0, sizeof(q.ind));
            std::memcpy(q.bxs[0], bx, sizeof(q.bxs[0]));
            q.ind[0][0] = p.ibeg; q.ind[0][1] = p.iend;
            q.nch = 1;
            sp.push_back(q);
        }
        if (sp.empty()) break;
        if (nodes[sp[0].node].level >= kMaxDepth)
            throw std::runtime_error("tree deeper than " + std::to_string(kMaxDepth) +
                                     " levels: more than maxparnode coincident centroids?");
        // x split (PARTITION_8, src/treecode3d_pb.f:315-325)
        {
            std::vector<Seg> segs;
            std::vector<int> who;
            for (size_t k = 0; k < sp.size(); ++k)
                if (sp[k].xl >= sp[k].critlen) {
                    segs.push_back({sp[k].ind[0][0], sp[k].ind[0][1], sp[k].xm});
                    who.push_back((int)k);
                }
            std::vector<int32_t> mid;
            pass(0, segs, mid);
            for (size_t w = 0; w < who.size(); ++w) {
                Split &q = sp[who[w]];
                q.ind[1][0] = mid[w] + 1; q.ind[1][1] = q.ind[0][1]; q.ind[0][1] = mid[w];
                std::memcpy(q.bxs[1], q.bxs[0], sizeof(q.bxs[0]));
                q.bxs[0][1] = q.xm; q.bxs[1][0] = q.xm;
                q.nch *= 2;
            }
        }
        // y split on each half (src/treecode3d_pb.f:327-339)
        {
            std::vector<Seg> segs;
            std::vector<std::pair<int, int>> who;
            for (size_t k = 0; k < sp.size(); ++k)
                if (sp[k].yl >= sp[k].critlen)
                    for (int i = 0; i < sp[k].nch; ++i) {
                        segs.push_back({sp[k].ind[i][0], sp[k].ind[i][1], sp[k].ym});
                        who.push_back({(int)k, i});
                    }
            std::vector<int32_t> mid;
            pass(1, segs, mid);
            for (size_t w = 0; w < who.size(); ++w) {
                Split &q = sp[who[w].first];
                const int i = who[w].second, nc = q.nch;
                q.ind[nc + i][0] = mid[w] + 1; q.ind[nc + i][1] = q.ind[i][1]; q.ind[i][1] = mid[w];
                std::memcpy(q.bxs[nc + i], q.bxs[i], sizeof(q.bxs[0]));
                q.bxs[i][3] = q.ym; q.bxs[nc + i][2] = q.ym;
            }
            for (size_t k = 0; k < sp.size(); ++k)
                if (sp[k].yl >= sp[k].critlen) sp[k].nch *= 2;
        }
        // z split on each quarter (src/treecode3d_pb.f:341-353)
        {
            std::vector<Seg> segs;
            std::vector<std::pair<int, int>> who;
            for (size_t k = 0; k < sp.size(); ++k)
                if (sp[k].zl >= sp[k].critlen)
                    for (int i = 0; i < sp[k].nch; ++i) {
                        segs.push_back({sp[k].ind[i][0], sp[k].ind[i][1], sp[k].zm});
                        who.push_back({(int)k, i});
                    }
            std::vector<int32_t> mid;
            pass(2, segs, mid);
            for (size_t w = 0; w < who.size(); ++w) {
                Split &q = sp[who[w].first];
                const int i = who[w].second, nc = q.nch;
                q.ind[nc + i][0] = mid[w] + 1; q.ind[nc + i][1] = q.ind[i][1]; q.ind[i][1] = mid[w];
                std::memcpy(q.bxs[nc + i], q.bxs[i], sizeof(q.bxs[0]));
                q.bxs[i][5] = q.zm; q.bxs[nc + i][4] = q.zm;
            }
            for (size_t k = 0; k < sp.size(); ++k)
                if (sp[k].zl >= sp[k].critlen) sp[k].nch *= 2;
        }
        // shrink boxes of children with >= 2 particles (src/treecode3d_pb.f:249-258)
        std::vector<int32_t> rng;
        std::vector<std::pair<int, int>> who;
        for (size_t k = 0; k < sp.size(); ++k)
            for (int i = 0; i < 8; ++i)
                if (sp[k].ind[i][0] < sp[k].ind[i][1]) {
                    rng.push_back(sp[k].ind[i][0]);
                    rng.push_back(sp[k].ind[i][1]);
                    who.push_back({(int)k, i});
                }
        std::vector<double> mm;
        minmax(rng, mm);
        for (size_t w = 0; w < who.size(); ++w)
            std::memcpy(sp[who[w].first].bxs[who[w].second], &mm[6 * w], 6 * sizeof(double));
        // children in creation order (src/treecode3d_pb.f:264-273)
        std::vector<int> next;
        for (auto &q : sp)
            for (int i = 0; i < q.nch; ++i)
                if (q.ind[i][0] <= q.ind[i][1]) {
                    LNode c;
                    c.ibeg = q.ind[i][0]; c.iend = q.ind[i][1];
                    c.level = nodes[q.node].level + 1; c.parent = q.node;
                    std::memcpy(c.bx, q.bxs[i], sizeof(c.bx));
                    nodes.push_back(c);
                    nodes[q.node].kids.push_back((int)nodes.size() - 1);
                    next.push_back((int)nodes.size() - 1);
                }
        level_nodes.swap(next);
    }
    if (cur != 0) {
        TCHK(hipMemcpyAsync(X[0], X[1], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        TCHK(hipMemcpyAsync(Y[0], Y[1], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        TCHK(hipMemcpyAsync(Z[0], Z[1], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        TCHK(hipMemcpyAsync(O[0], O[1], sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    // depth-first pre-order numbering, as the recursive build creates nodes
    t = HostTree();
    t.n = n;
    t.order.resize(n);
    TCHK(hipMemcpyAsync(t.order.data(), O[0], sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    std::vector<int> dfs, newid(nodes.size());
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int id = stack.back();
        stack.pop_back();
        newid[id] = (int)dfs.size();
        dfs.push_back(id);
        const auto &k = nodes[id].kids;
        for (int i = (int)k.size() - 1; i >= 0; --i) stack.push_back(k[i]);
    }
    for (int id : dfs) {
        const LNode &p = nodes[id];
        const double xm = (p.bx[1] + p.bx[0]) / 2.0, ym = (p.bx[3] + p.bx[2]) / 2.0,
                     zm = (p.bx[5] + p.bx[4]) / 2.0;
        const double t1 = p.bx[1] - xm, t2 = p.bx[3] - ym, t3 = p.bx[5] - zm;
        t.ibeg.push_back(p.ibeg); t.iend.push_back(p.iend); t.numpar.push_back(p.iend - p.ibeg + 1);
        t.level.push_back(p.level); t.nchild.push_back((int)p.kids.size());
        t.child_off.push_back((int)t.children.size());
        for (int k : p.kids) t.children.push_back(newid[k]);
        t.xmin.push_back(p.bx[0]); t.xmax.push_back(p.bx[1]); t.ymin.push_back(p.bx[2]);
        t.ymax.push_back(p.bx[3]); t.zmin.push_back(p.bx[4]); t.zmax.push_back(p.bx[5]);
        t.xmid.push_back(xm); t.ymid.push_back(ym); t.zmid.push_back(zm);
        t.radius.push_back(std::sqrt(t1 * t1 + t2 * t2 + t3 * t3));
        if (p.kids.empty()) t.leaves.push_back(newid[id]);
    }
    TCHK(hipStreamSynchronize(s));
}

}  // namespace tabi
